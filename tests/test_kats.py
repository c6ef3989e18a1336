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


@pytest.mark.gpu
@pytest.mark.parametrize("name", K.NAMES)
def test_kat_gpu(gpu, name):
    print(f"[gpu kat] {name}", flush=True)
    _check(getattr(K, "kat_" + name)(gpu, K.KATS[name]))
