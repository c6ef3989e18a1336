"""Reference known-answer tests (CUDA/test/*.cc EXPECT_NEAR values) run on the
FP64 oracle (CPU) and on the HIP library (GPU)."""
import pytest

import kat_cases as K


def _check(results):
    bad = [(lbl, got, exp, tol) for lbl, got, exp, tol in results if not abs(got - exp) <= tol]
    assert not bad, "\n".join(f"{l}: got {g!r} expected {e!r} (tol {t})" for l, g, e, t in bad)


@pytest.mark.parametrize("name", K.NAMES)
def test_kat_oracle(cpu, name):
    _check(getattr(K, "kat_" + name)(cpu, K.KATS[name]))


# Whole-loop KATs (500 iterations): the reference's values are float32 results
# printed to three decimals, and at least one sits at the edge of its own EPS
# from the exact value (ictgv2.cc:676 expects 184.705; the FP64 oracle gives
# 184.705999).  On the GPU these are checked two ways: against the FP64 oracle
# (which passes the KAT at EPS) within 2e-6 of the largest checked magnitude,
# float32 rounding over 500 iterations, and against the KAT value within EPS
# plus that same margin.
LOOP_KATS = {"tv_iteration", "tgv_iteration", "ictgv_iteration"}
F32_LOOP_REL = 2e-6


@pytest.mark.gpu
@pytest.mark.parametrize("name", K.NAMES)
def test_kat_gpu(gpu, cpu, name):
    print(f"[gpu kat] {name}", flush=True)
    got = getattr(K, "kat_" + name)(gpu, K.KATS[name])
    if name in LOOP_KATS:
        ref = getattr(K, "kat_" + name)(cpu, K.KATS[name])
        margin = F32_LOOP_REL * max(abs(o) for _, o, _, _ in ref)
        for (lbl, g, e, t), (_, o, _, _) in zip(got, ref):
            assert abs(g - o) <= margin, f"{lbl}: GPU {g!r} vs FP64 oracle {o!r}"
        got = [(lbl, g, e, t + margin) for lbl, g, e, t in got]
    _check(got)
