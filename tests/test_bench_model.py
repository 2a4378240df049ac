"""The MFMA flop model that bench.py uses for its MFMA rates, against the
executed BF16 MFMA flops that the hardware counted (rocprofv3 MfmaFlopsBF16 =
512 x SQ_INSTS_VALU_MFMA_MOPS_BF16 per dispatch, MI355X).

Counted values per dispatch:
  k_inv<127, 3, INV_SCALAR, true, 2> (sp2fc, 121 field-levels, 8 members): 93.62 GFLOP
  k_inv_rt<42, 3, 2> (T42 L20 round-trip inverse, 128 members):           51.64 GFLOP
  k_inv<127, 3, INV_UV, true, 1> (dv2uv, 40 levels, 8 members):           241.59 GFLOP
"""
import importlib.util
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_mfma_flop_model_matches_counters():
    b = _bench()
    f = b._mfma_flops_inverse
    # T127 sp2fc: 4 fields per block (tile variants 0 and 2)
    assert abs(f(127, 121, 8, 4, 1) / 93.62e9 - 1) < 1e-3
    # T42 fused round-trip inverse: qi (61 fields) + qj (1 field) with 8 fields
    # per block, dv2uv (20 levels, 4 per block, two terms)
    rt = f(42, 61, 128, 8, 1) + f(42, 1, 128, 8, 1) + f(42, 20, 128, 4, 2)
    assert abs(rt / 51.64e9 - 1) < 1e-3
    # T127 dv2uv with tile variant 1: 2 fields per block (1 level), two terms
    assert abs(f(127, 40, 8, 1, 2) / 241.59e9 - 1) < 1e-3
