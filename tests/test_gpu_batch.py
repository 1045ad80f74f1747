"""Batches of frame models under the nonlinear Newmark driver (main.c:3305-3960),
side by side on the GPU, through the C ABI (cb_batch_*).

- A mixed batch: towers and lattice girders of 60-300 members, without and
  with numerical dissipation, girders whose band has two tiles below the
  diagonal (factored by a pair of workgroups), and models with itemax 1, whose Newton loops
  fail and halve the time step (main.c:3826-3846) while the other models of
  the batch step on.  Every model's path is the path of the same model run
  alone (cb_run): identical iteration counts, displacements to 1e-12
  relative; and the oracle's path to 1e-9 with identical iteration counts.
- Girders with three tiles below the diagonal, factored through memory
  rather than in LDS, against their single runs and the oracle.
- The 512-frame workload (200-5000 members each, about 2.66 million
  equations in all) for 20 time steps: every model runs to the step limit
  with finite displacements, and the largest and smallest models and the
  largest model with a half-bandwidth above 64 match their single runs to
  1e-12 with identical iteration counts.
"""
import os
import tempfile

import numpy as np
import pytest

import cubens
import oracle as O
from cubens import inputs
from helpers import model_file, relerr

pytestmark = pytest.mark.gpu

# the last two are girders with a half-bandwidth above 64 (two tiles below the
# diagonal), which the batch factors with a pair of workgroups
MIXED = [
    dict(n=60, kind="tower", amp=8e4),
    dict(n=150, kind="grid", amp=8e4, itemax=1),
    dict(n=300, kind="tower", amp=3e5, numopt=1, spectrds=0.8),
    dict(n=90, kind="grid", amp=1e6),
    dict(n=200, kind="tower", amp=1e6, itemax=1),
    dict(n=120, kind="grid", amp=3e5, numopt=2, spectrds=0.7),
    dict(n=150, kind="grid", amp=3e5, instance=114),
    dict(n=150, kind="grid", amp=8e4, itemax=1, instance=104),
]


def _mixed_files(steps):
    paths = []
    for i, kw in enumerate(MIXED):
        kw = dict(kw)
        n = kw.pop("n")
        inst = kw.pop("instance", 100 + i)
        txt = inputs.space_frame(n, instance=inst, steps=steps, **kw).text
        paths.append(model_file(f"batch_{i}.txt", txt))
    return paths


def _single(path, steps, sweep=cubens.SWEEP_CHAIN):
    with cubens.GpuModel(path) as g:
        r = g.run(cubens.options(cubens.SOLVER_BAND64, max_steps=steps, sweep_schedule=sweep))
    assert r["status"] == 0, r["msg"]
    return r


def test_mixed_batch_matches_single_runs_and_oracle():
    steps = 15
    paths = _mixed_files(steps)
    with cubens.GpuBatch(paths) as b:
        st = b.run(max_steps=steps)
        assert st["failed"] == 0 and np.all(st["status"] == 0), st
        assert np.all(st["nsteps"] == steps)
        halving = 0
        for i, p in enumerate(paths):
            rb = b.path(i)
            rs = _single(p, steps)
            ro = O.OracleModel(p).run(steps)
            assert ro["rc"] == 0, ro["msg"]
            halving += ro["n_newton"] > ro["itecnt"].sum()
            print(f"model {i}: neq {b.info[i].neq}, Newton {ro['n_newton']}, "
                  f"batch vs single {relerr(rb['d'], rs['d']):.1e}, vs oracle {relerr(rb['d'], ro['d']):.1e}")
            assert np.array_equal(rb["itecnt"], rs["itecnt"])
            assert np.array_equal(rb["lpf"], rs["lpf"])
            assert relerr(rb["d"], rs["d"]) < 1e-12
            assert np.array_equal(rb["itecnt"], ro["itecnt"])
            assert relerr(rb["d"], ro["d"]) < 1e-9
        assert halving >= 2, "the batch should hold models that halve their time step"
        assert sum(inf.bandwidth > 64 for inf in b.info) >= 2, "the batch should hold two-tile-band models"
        assert st["newton"] == sum(O.OracleModel(p).run(steps)["n_newton"] for p in paths)


def test_batch_rejects_models_it_cannot_run():
    shell = model_file("batch_shell.txt", inputs.shell_panel(n=4).text)
    tower = model_file("batch_tower.txt", inputs.space_frame(40, kind="tower", steps=5).text)
    with pytest.raises(cubens.CubensError, match="frame-only|Newmark"):
        cubens.GpuBatch([tower, shell])


def test_frame_workload_batch_512_models():
    steps = 20
    d = tempfile.mkdtemp(prefix="cubens_frames_")
    models = inputs.frame_batch(steps=steps)
    paths = []
    for i, m in enumerate(models):
        p = os.path.join(d, f"f{i:03d}.txt")
        with open(p, "w") as fh:
            fh.write(m.text)
        paths.append(p)
    with cubens.GpuBatch(paths) as b:
        neq = np.array([inf.neq for inf in b.info])
        nfr = np.array([inf.ne_fr for inf in b.info])
        assert b.count == 512 and nfr.min() >= 200 and nfr.max() <= 5000
        st = b.run(max_steps=steps)
        print(f"512 frames, {nfr.sum()} members, {neq.sum()} equations: {st['rounds']} rounds, "
              f"{st['newton']} Newton iterations in {st['t_total']:.2f} s")
        assert st["failed"] == 0 and np.all(st["nsteps"] == steps), st
        bw = np.array([inf.bandwidth for inf in b.info])
        wide = np.flatnonzero(bw > 64)
        for i in (int(np.argmax(neq)), int(np.argmin(neq)), int(wide[np.argmax(neq[wide])])):
            rb = b.path(i)
            assert np.all(np.isfinite(rb["d"]))
            rs = _single(paths[i], steps)
            assert np.array_equal(rb["itecnt"], rs["itecnt"])
            assert relerr(rb["d"], rs["d"]) < 1e-12
        for i in range(0, 512, 37):
            assert np.all(np.isfinite(b.path(i)["d"]))


def test_batch_wide_band_models():
    """Girders whose band has three tiles below the diagonal: their workgroup
    factors them with the panels and updates through memory (not in LDS).
    Batch paths equal the single runs and the oracle's."""
    steps = 10
    specs = [dict(n=300, shape=(5, 5), inst=131, amp=3e5), dict(n=400, shape=(6, 5), inst=132, amp=8e4),
             dict(n=120, shape=None, inst=133, amp=1e6)]
    paths = []
    for i, sp in enumerate(specs):
        kind = "grid" if sp["shape"] else "tower"
        txt = inputs.space_frame(sp["n"], kind=kind, instance=sp["inst"], steps=steps, amp=sp["amp"],
                                 slice_shape=sp["shape"]).text
        paths.append(model_file(f"batch_wide_{i}.txt", txt))
    with cubens.GpuBatch(paths) as b:
        assert sum(inf.bandwidth > 128 for inf in b.info) == 2, [inf.bandwidth for inf in b.info]
        st = b.run(max_steps=steps)
        assert st["failed"] == 0 and np.all(st["status"] == 0) and np.all(st["nsteps"] == steps), st
        for i, p in enumerate(paths):
            rb = b.path(i)
            rs = _single(p, steps)
            ro = O.OracleModel(p).run(steps)
            assert ro["rc"] == 0, ro["msg"]
            print(f"model {i}: neq {b.info[i].neq}, bandwidth {b.info[i].bandwidth}, "
                  f"batch vs single {relerr(rb['d'], rs['d']):.1e}, vs oracle {relerr(rb['d'], ro['d']):.1e}")
            assert np.array_equal(rb["itecnt"], rs["itecnt"])
            assert relerr(rb["d"], rs["d"]) < 1e-12
            assert np.array_equal(rb["itecnt"], ro["itecnt"])
            assert relerr(rb["d"], ro["d"]) < 1e-9


def test_batch_edge_cases():
    """A batch of one model equals its single run bitwise (with the single
    run's sweeps in the two-product row order, the order the batch's solve
    keeps; the default chain-matrix sweeps differ in rounding); zero steps record nothing;
    a run without the displacement path refuses to return one; a second run
    of the same batch repeats the first bitwise."""
    p = model_file("batch_one.txt", inputs.space_frame(80, kind="tower", instance=7, steps=8, amp=1e6).text)
    with cubens.GpuBatch([p]) as b:
        st = b.run(max_steps=0)
        assert st["rounds"] == 0 and np.all(st["nsteps"] == 0)
        st = b.run(max_steps=8)
        assert np.all(st["status"] == 0) and np.all(st["nsteps"] == 8)
        first = b.path(0)
        rs = _single(p, 8, sweep=cubens.SWEEP_ROWS)
        assert np.array_equal(first["itecnt"], rs["itecnt"]) and np.array_equal(first["d"], rs["d"])
        again = b.run(max_steps=8)
        assert again["newton"] == st["newton"]
        assert np.array_equal(b.path(0)["d"], first["d"])
        b.run(max_steps=4, keep_d=False)
        with pytest.raises(cubens.CubensError, match="path"):
            b.path(0)
        with pytest.raises(cubens.CubensError, match="range"):
            lib = cubens.lib()
            import ctypes as C
            cubens._check(lib.cb_batch_get_path(b.h, 5, None, None, None))


def test_graph_replay_equals_the_launches_one_by_one():
    """The round's Newton iteration captured once and replayed as a HIP graph
    (no memset node in it; the band, flags and counters are cleared by
    kernels) gives every model's path bitwise as the launches issued one by
    one, over every round of a mixed batch that halves time steps, and again
    on a second graph run of the same batch, with eager runs in between."""
    steps = 12
    paths = _mixed_files(steps)
    with cubens.GpuBatch(paths) as b:
        eager = b.run(max_steps=steps)
        ref = [b.path(i) for i in range(b.count)]
        for rep in range(2):
            print(f"[batch] graph run {rep}", flush=True)
            st = b.run(max_steps=steps, graph=True)
            assert np.array_equal(st["status"], eager["status"])
            assert np.array_equal(st["nsteps"], eager["nsteps"])
            assert st["rounds"] == eager["rounds"]
            for i in range(b.count):
                p = b.path(i)
                assert np.array_equal(p["itecnt"], ref[i]["itecnt"])
                assert np.array_equal(p["lpf"], ref[i]["lpf"])
                assert np.array_equal(p["d"], ref[i]["d"])
            print(f"[batch] ops per round: eager {eager['ops_per_round']:.1f}, graph "
                  f"{st['ops_per_round']:.1f}; iteration {st['body_ops']} operations", flush=True)
            assert st["body_ops"] == eager["body_ops"] > 10
            assert st["ops_per_round"] < eager["ops_per_round"] - st["body_ops"] + 2
            b.run(max_steps=3)                 # eager work between the graph runs


def test_autotuned_helper_pairs_keep_every_path(tmp_path, monkeypatch):
    """The tuner times the band factor for 0, 1 and 2 helper pairs (the two
    girders with two tiles below the diagonal), keeps the fastest, and every
    model's path stays bitwise the same, launched one by one and replayed as a
    graph captured after the tuning; the choice goes to CUBENS_TUNE_FILE and a
    second batch of the same shape takes it without timing again."""
    steps = 8
    paths = _mixed_files(steps)
    tune = tmp_path / "tune.txt"
    monkeypatch.setenv("CUBENS_TUNE_FILE", str(tune))
    with cubens.GpuBatch(paths) as b:
        plain = b.run(max_steps=steps)
        assert plain["pairs"] == plain["pairs_max"] == 2 and plain["tune_s"] == 0
        ref = [b.path(i) for i in range(b.count)]
        print("[batch] autotune", flush=True)
        st = b.run(max_steps=steps, autotune=True)
        trials = b.tuning()
        print(f"[batch] tuner trials (pairs, ms): {trials}; kept {st['pairs']} in {st['tune_s']:.3f}s", flush=True)
        assert [p for p, _ in trials] == [0, 1, 2]
        assert all(ms > 0 for _, ms in trials)
        assert st["pairs"] == min(trials, key=lambda t: t[1])[0] and st["tune_s"] > 0
        for run in (st, b.run(max_steps=steps, graph=True)):
            assert np.array_equal(run["nsteps"], plain["nsteps"]) and run["rounds"] == plain["rounds"]
            for i in range(b.count):
                p = b.path(i)
                assert np.array_equal(p["itecnt"], ref[i]["itecnt"])
                assert np.array_equal(p["d"], ref[i]["d"])
        kept = st["pairs"]
    lines = tune.read_text().splitlines()
    assert len(lines) == 1 and lines[0].startswith("batch-pairs ") and lines[0].endswith(f" {kept}")
    with cubens.GpuBatch(paths) as b2:
        st2 = b2.run(max_steps=2, autotune=True)
        assert b2.tuning() == [] and st2["pairs"] == kept


def test_fp32_element_tangents_follow_the_fp64_paths():
    """The member tangents in FP32 (rounded once into the FP64 band; internal
    forces, residual, factor and solve FP64): every model runs its steps with
    the FP64 batch's step count and times, Newton iterations within 10%, and
    the models whose Newton loops converge keep the FP64 path's displacements
    to 1e-9 (measured 2e-12; the error is printed), launched one by one and as
    a graph captured for the FP32 tangent.  The models with itemax 1 accept
    the state after a single correction, which follows the tangent itself:
    for them only the steps and times are compared."""
    steps = 8
    paths = _mixed_files(steps)
    with cubens.GpuBatch(paths) as b:
        ref_st = b.run(max_steps=steps)
        ref = [b.path(i) for i in range(b.count)]
        for graph in (False, True):
            st = b.run(max_steps=steps, fp32_elements=True, graph=graph)
            assert np.array_equal(st["status"], ref_st["status"])
            assert np.array_equal(st["nsteps"], ref_st["nsteps"])
            errs = []
            for i in range(b.count):
                p = b.path(i)
                assert np.array_equal(p["lpf"], ref[i]["lpf"])          # the step times
                ni, nr = int(p["itecnt"].sum()), int(ref[i]["itecnt"].sum())
                assert abs(ni - nr) <= max(1, nr // 10), (i, ni, nr)
                if MIXED[i].get("itemax", 20) > 1:
                    errs.append(float(np.abs(p["d"] - ref[i]["d"]).max() / np.abs(ref[i]["d"]).max()))
            print(f"[batch] fp32 elements (graph {graph}): Newton {st['newton']} (FP64 {ref_st['newton']}), "
                  f"worst displacement error of the converging models {max(errs):.2e}", flush=True)
            assert len(errs) == 5 and max(errs) < 1e-9
        # and back to the FP64 tangent: the FP64 paths bitwise
        b.run(max_steps=steps, graph=True)
        for i in range(b.count):
            assert np.array_equal(b.path(i)["d"], ref[i]["d"])
