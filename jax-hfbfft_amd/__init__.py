"""MI355X-native static Skyrme-HF(+BCS) iteration of ascsn/jax-hfbfft.

The package directory is ``jax-hfbfft_amd``; import it as ``jax_hfbfft_amd``
via ``__graft_entry__.load_package()``.  Submodules mirror the reference's
module names (static, meanfield, densities, coulomb, levels, ...) and run on
hand-written HIP kernels in ``libhfbhip.so`` (C ABI in include/hfb_hip.h).
"""
import importlib as _importlib

_SUBMODULES = ("_lib", "kernels", "adjoint", "driver", "static", "meanfield", "densities", "coulomb", "levels", "grids",
               "forces", "params", "energies", "inout", "pairs", "trivial", "moment", "output", "reader", "constraint")


def __getattr__(name):
    if name in _SUBMODULES:
        return _importlib.import_module(f"{__name__}.{name}")
    raise AttributeError(name)
