"""GPU parity of the wave-cooperative ERI kernel (the engine of every class
with total angular momentum L >= 5, and of all classes when
RCHEM_ERI_GENERIC=1 is set before a context is created).

Bars: integrals within 1e-12 absolute of the reference's contr_coulomb
golden values (compat Boys); J and K from the direct build within 1e-11 of
J and K contracted from the same library's ERI tensor; tensor entries within
1e-12 of the oracle's contracted THO integrals (exact Boys)."""
import os

import numpy as np
import pytest

import oracle as O
from rchem_amd import basis as rb
from rchem_amd._lib import BOYS_COMPAT, BOYS_EXACT, default_opts
from rchem_amd.inputs import synthetic_shells

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def generic_ctx():
    """A context whose every ERI class runs on the wave-cooperative kernel."""
    os.environ["RCHEM_ERI_GENERIC"] = "1"
    try:
        ctx = rb.Context(0)
    finally:
        del os.environ["RCHEM_ERI_GENERIC"]
    yield ctx
    ctx.close()


def test_generic_contracted_samples_all_classes(generic_ctx):
    """All 21 classes through the wave-cooperative kernel against the
    reference's contr_coulomb values, in the quartets' own shell order."""
    g = np.load(os.path.join(GOLD, "contracted_samples.npz"))
    _say(f"case: generic kernel, {len(g['quartets'])} contracted samples")
    B = rb.Basis.from_shells(g["centres"], g["ls"], g["nprim"], g["exps"], g["coefs"],
                             ctx=generic_ctx)
    blocks = rb.eri_shell_quartets(B, g["quartets"],
                                   default_opts(boys_mode=BOYS_COMPAT, schwarz_thresh=0.0))
    qof = g["quartet_of"]
    worst = max(np.max(np.abs(blk.reshape(-1) - g["values"][qof == qi]))
                for qi, blk in enumerate(blocks))
    assert worst < 1e-12


def _dshell_basis(ctx, nshell=10, seed=7, box=5.0):
    c, l, npr, ex, cf = synthetic_shells(nshell=nshell, seed=seed, box=box)
    l[:3] = 2  # at least three d shells, so every L >= 5 class with two of them appears
    return rb.Basis.from_shells(c, l, npr, ex, cf, ctx=ctx), (c, l, npr, ex, cf)


@pytest.mark.parametrize("generic", [False, True])
def test_dshell_jk_matches_tensor(gpu_ctx, generic_ctx, generic):
    """Direct J/K (screened at 1e-14) against J/K contracted from the ERI
    tensor, on a 10-shell basis with d shells; the tensor's entries against the
    oracle at sampled positions."""
    ctx = generic_ctx if generic else gpu_ctx
    B, sh = _dshell_basis(ctx)
    n = B.nbf
    _say(f"case: d-shell basis nbf={n}, generic={generic}")
    o = default_opts(boys_mode=BOYS_EXACT, schwarz_thresh=0.0)
    I = rb.build_I(B, o)
    rng = np.random.default_rng(3)
    X = rng.standard_normal((n, n))
    D = 0.05 * (X + X.T)
    J = np.zeros((n, n)); K = np.zeros((n, n))
    rb.JK_direct(J, K, B, D, default_opts(boys_mode=BOYS_EXACT, schwarz_thresh=1e-14))
    Jt, Kt = rb.JK_inmem(I, D)
    assert np.max(np.abs(J - Jt)) < 1e-11
    assert np.max(np.abs(K - Kt)) < 1e-11
    # 8-fold symmetry of the tensor (a quartet whose bra and ket are the same
    # shell pair evaluates (ab|cd) and (cd|ab) by different recurrence paths:
    # equal to rounding, 1e-13 relative to the largest integral) and sampled
    # entries against the oracle
    assert np.max(np.abs(I - I.transpose(1, 0, 2, 3))) == 0.0
    assert np.max(np.abs(I - I.transpose(2, 3, 0, 1))) < 1e-13 * np.max(np.abs(I))
    Bo = O.cgtos_from_shells(*sh)
    idx = rng.integers(0, n, (60, 4))
    ref = np.array([O.contracted_eri(Bo, *q, boys=O.BOYS_EXACT) for q in idx])
    got = np.array([I[tuple(q)] for q in idx])
    assert np.max(np.abs(got - ref)) < 1e-12


def test_dshell_jk_mixed_precision(gpu_ctx):
    """The FP32 kernels of the d classes (per-lane, LDS digestion up to L = 4):
    a mixed build with a loose threshold (many FP32 quartets) stays within
    FP32 rounding of the all-FP64 build."""
    from rchem_amd._lib import PREC_MIXED
    B, _ = _dshell_basis(gpu_ctx)
    n = B.nbf
    rng = np.random.default_rng(5)
    X = rng.standard_normal((n, n))
    D = 0.05 * (X + X.T)
    _say(f"case: d-shell basis nbf={n}, mixed precision (threshold 1e-3) vs FP64")
    J = np.zeros((n, n)); K = np.zeros((n, n))
    rb.JK_direct(J, K, B, D, default_opts(schwarz_thresh=1e-14))
    Jm = np.zeros((n, n)); Km = np.zeros((n, n))
    rb.JK_direct(Jm, Km, B, D, default_opts(schwarz_thresh=1e-14, precision=PREC_MIXED,
                                            mixed_thresh=1e-3))
    st = rb.jk_last_stats(B)
    assert st.quartets_fp32 > 0
    scale = max(np.max(np.abs(J)), np.max(np.abs(K)))
    assert np.max(np.abs(Jm - J)) < 1e-6 * scale
    assert np.max(np.abs(Km - K)) < 1e-6 * scale
