"""Norms (trivial.py:5-12), computed on the device in FP64."""
import torch


def rpsnorm(ps: torch.Tensor, wxyz: float):
    return wxyz * torch.sum((ps.conj() * ps).real.to(torch.float64))


def overlap(pl: torch.Tensor, pr: torch.Tensor, wxyz: float):
    return wxyz * torch.sum((pl.conj() * pr).to(torch.complex128))
