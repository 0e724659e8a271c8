"""cusnarks_amd -- MI355X-native Groth16 prover core (BN254).

Hand-written HIP kernels for gfx950 behind a C ABI (include/csk.h, built into
cusnarks_amd/libcsk.so), with a Python host layer that mirrors cusnarks'
pycusnarks binding (``cusnarks_amd.pycusnarks``) so existing callers of the
GPU multi-exponentiation and polynomial path work unchanged.
"""
__version__ = "0.1.0"


def build(force: bool = True) -> str:
    from .build import build as _b
    return _b(force=force)
