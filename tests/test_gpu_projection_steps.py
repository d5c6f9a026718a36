"""Time-stepped projection with the operator and hierarchy reused across steps and the
V-cycles replayed from captured graphs: after each correction the divergence D u + flux equals
the pressure solve's residual (below 1e-8 |b|), the mixed-precision and FP64 V-cycle pressures
agree to the solve tolerance, and the final adjoint solve converges."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_projection_steps(lsx):
    import lsx_pkg
    out = lsx_pkg.projection.run(dims=(96, 64, 64), steps=3, graphs=True, compare_fp64=True, adjoint=True)
    for r in out["records"]:
        print({k: (round(v, 6) if isinstance(v, float) and k.endswith("_s") else v) for k, v in r.items()}, flush=True)
        assert r["divergence_after"] <= 1.05e-8 * r["rhs_norm"]
        assert r["divergence_before"] > 1e3 * r["divergence_after"]
        assert r["dp_rel_diff_vs_fp64"] <= 1e-4
        assert r["iterations"] <= r["fp64_iterations"] + 2
    print(out["adjoint"], flush=True)
    assert out["adjoint"]["residual_rel"] <= 1e-8
    assert out["adjoint"]["iterations"] < 60
