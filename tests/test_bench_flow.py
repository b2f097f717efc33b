"""bench.py's whole flow on the CPU with a stand-in for the device library:
the driver's JSON line (metric, value, ms_per_step, roofline with the split
near-field kernel's figures, cpu_baseline switched off, the reference
geometry's entry) comes out with its keys and arithmetic intact, so a change
to the bench's own plumbing is caught without a GPU.  The stand-in returns
fixed per-step statistics; nothing here measures anything."""
import io
import json
import os
import sys
import types
from contextlib import redirect_stdout

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def stub_library(split=128):
    m = types.ModuleType("isg")

    class IsgError(RuntimeError):
        pass

    class Solver:
        KERNELS = {"near": 0, "m2l": 1, "bicgs_update": 2, "near_fp32": 3, "m2l_fp32": 4, "p2m": 5, "interp": 6,
                   "receivers": 7}

        def __init__(self, max_rhs=1, level=4, numrx=16, numtx=16, precision=0, **kw):
            self.level, self.max_rhs, self.precision = level, max_rhs, precision
            self.numsamp = [4 * (8 + i) for i in range(level)]
            self.N = (8 << (level - 1)) ** 2
            self.timing = False

        def geometry(self):
            return np.zeros(self.N, complex), np.arange(self.N)

        def close(self):
            pass

        def set_progress(self, fn=None):
            if fn:
                fn(self.max_rhs, self.max_rhs)

        def profile(self, enable=True):
            self.timing = enable

        def counters(self):
            return {"refinements": 0}

        def profile_read(self, kernel):
            # 10 launches of 1 ms per timed region for every group
            return (10.0, 10) if self.timing else (0.0, 0)

        def tuning_get(self, precision, batch_class, stage, level=0):
            if stage == "near":
                return split or 32
            return 1 if stage == "leaf" else -1

    class Dbim:
        def __init__(self, solver, otrue, field_start=0, measurements=None):
            self.s, self.n = solver, solver.N
            self.setup_stats = {"solves": 4, "solve_iters": 40, "not_converged": 0}

        def get(self, name):
            return np.ones(self.n if name != "mes" else 16, complex)

        def iterate(self):
            batched = self.s.max_rhs > 1
            return {"seconds": 0.01, "solves": 48, "solve_iters": 100, "field_iters": 30, "not_converged": 0,
                    "alpha": 1e-3, "misfit": 0.5, "matvecs": 300, "matvecs_fp32": 250 if batched else 0,
                    "net_seconds": 0.002, "lambda_abs": 0.9}

        def restart(self):
            pass

        def set_warmnet(self, net):
            pass

    class WarmNet:
        def __init__(self, **kw):
            pass

    m.IsgError, m.Solver, m.Dbim, m.WarmNet = IsgError, Solver, Dbim, WarmNet
    m.raster_to_tree = lambda raster, r2t: np.asarray(raster).ravel()[r2t]
    m.phantom_ellipses = lambda side, n, seed: np.zeros((side, side), complex)
    return m


def run_bench(monkeypatch, *args, split=128):
    monkeypatch.setitem(sys.modules, "isg", stub_library(split))
    monkeypatch.syspath_prepend(ROOT)
    import bench
    monkeypatch.setattr(sys, "argv", ["bench.py", *args])
    out = io.StringIO()
    with redirect_stdout(out):
        bench.main()
    lines = [ln for ln in out.getvalue().splitlines() if ln.strip()]
    assert len(lines) == 1, lines
    return json.loads(lines[0])


def test_bench_json_line_with_stub_library(monkeypatch):
    d = run_bench(monkeypatch, "--workload", "dbim-256", "--steps", "2", "--warmup", "1", "--cpu-baseline", "off")
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
              "vs_baseline", "dtype", "data", "config", "roofline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 2 and d["warmup"] == 1 and d["higher_is_better"] is False
    assert "workload" in d["config"]
    assert abs(d["ms_per_step"] - 1e3 * d["value"]) < 0.1
    r = d["roofline"]
    assert r["kernel"].startswith("k_near_h16") and r["peak"] == 2500.0
    assert abs(r["frac"] - r["achieved"] / r["peak"]) < 1e-3
    assert abs(r["frac_of_split_ceiling"] - 3 * r["frac"]) < 1e-3
    # 2 timed steps x 250 FP32 matvecs x the level-6 near-field flops / 10 ms of near-field time
    level = 6
    nb = 1 << (level - 1)
    flops = 8 * 64 * 64 * (3 * nb - 2) ** 2 + 8 * 64 * (4 * (8 + level - 1)) * nb * nb
    assert abs(r["achieved"] - 2 * 250 * flops / 10e-3 / 1e12) < 0.01
    assert d["speedup_vs_plain"] > 0 and "fast" in d and "plain" in d
    assert "reference_geometry" in d and "learned_s_per_iter" in d["reference_geometry"]
    # warm-up 1 + 2 timed steps of the 5-iteration cycle end on iteration 3; the reference
    # geometry's timed cycle ends on the cycle's last iteration
    assert d["contrast_after_iteration"] == 3
    assert d["reference_geometry"]["contrast_after_iteration"] == 5
    assert "cpu_baseline" not in d


def test_bench_roofline_fp32_kernel_when_not_split(monkeypatch):
    d = run_bench(monkeypatch, "--workload", "dbim-256", "--steps", "1", "--warmup", "0", "--cpu-baseline", "off",
                  "--ref-geometry", "off", split=0)
    r = d["roofline"]
    assert r["kernel"].startswith("k_near_mfma") and r["peak"] == 157.3
    assert r["frac_of_split_ceiling"] is None
    assert "reference_geometry" not in d


def test_bench_roofline_wide_split_kernel(monkeypatch):
    d = run_bench(monkeypatch, "--workload", "dbim-256", "--steps", "1", "--warmup", "0", "--cpu-baseline", "off",
                  "--ref-geometry", "off", split=256)
    r = d["roofline"]
    assert r["kernel"].startswith("k_near_h16w") and r["peak"] == 2500.0
    assert abs(r["frac_of_split_ceiling"] - 3 * r["frac"]) < 1e-3
