"""The C-ABI library loads and exports every symbol include/cravagpu.h declares (no GPU calls)."""
import ctypes

import crava_amd as C


def test_library_exports_every_declared_symbol():
    syms = C.header_symbols()
    assert len(syms) >= 25
    L = ctypes.CDLL(C.LIB_PATH)
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_host_setup_entry_points_without_gpu():
    # cg_factorable and cg_build_kz_tables are host-only
    assert C.lib().cg_factorable(512 + 64) == 576
    assert C.lib().cg_factorable(577) == 588
    assert C.lib().cg_factorable(1) == 1


def test_bad_arguments_are_rejected_without_gpu():
    # argument checks come before any device call
    assert C.lib().cg_ctx_create(None, 0, 0) == C.CG_EINVAL
    assert C.lib().cg_grid_destroy(None) == C.CG_EINVAL
    assert C.lib().cg_posterior(None, None) == C.CG_EINVAL
    assert C.lib().cg_simulate(None, None) == C.CG_EINVAL
    assert C.lib().cg_last_error(None) == b"null context"
