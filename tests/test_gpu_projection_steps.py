"""Time-stepped projection with the operator and hierarchy reused across steps and the
V-cycles replayed from captured graphs: after each correction the divergence D u + flux equals
the pressure solve's residual (below 1e-8 |b|), the mixed-precision and FP64 V-cycle pressures
agree to the solve tolerance, and the final adjoint solve converges."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,steps", [((96, 64, 64), 3), ((384, 256, 256), 20)], ids=["96x64x64", "c5_384x256x256"])
@pytest.mark.timeout(900)
def test_projection_steps(lsx, dims, steps):
    """The second case is C5 at its size: 384x256x256 cells, sphere r = 0.1, dt = 0.002, 20 steps
    on one operator and hierarchy with graph-replayed V-cycles, then the adjoint of the drag."""
    import lsx_pkg

    def log(r):
        print(f"dims={dims} step {r['step']}: it {r['iterations']} ({r['solve_s'] * 1e3:.1f} ms) fp64 it "
              f"{r['fp64_iterations']} ({r['fp64_solve_s'] * 1e3:.1f} ms) divergence {r['divergence_before']:.3e} -> "
              f"{r['divergence_after']:.3e} dp diff {r['dp_rel_diff_vs_fp64']:.2e}", flush=True)

    out = lsx_pkg.projection.run(dims=dims, steps=steps, graphs=True, compare_fp64=True, adjoint=True, log=log)
    assert len(out["records"]) == steps
    for r in out["records"]:
        print({k: (round(v, 6) if isinstance(v, float) and k.endswith("_s") else v) for k, v in r.items()}, flush=True)
        assert r["divergence_after"] <= 1.05e-8 * r["rhs_norm"]
        assert r["divergence_before"] > 1e3 * r["divergence_after"]
        assert r["dp_rel_diff_vs_fp64"] <= 1e-4
        assert r["iterations"] <= r["fp64_iterations"] + 2
    print(out["adjoint"], flush=True)
    assert out["adjoint"]["residual_rel"] <= 1e-8
    assert out["adjoint"]["iterations"] < 60
