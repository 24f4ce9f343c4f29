"""YAML configuration reader (reader.py of the reference); safe loader only."""
import yaml


def read_yaml(path):
    with open(path) as f:
        return yaml.safe_load(f)
