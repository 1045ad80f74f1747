"""Writes the golden checkpoint fixtures of tests/golden/checkpoint/ with the
reference's own checkpoint reader and writer, restartStep and checkPoint
(misc.c:494-716), compiled from its sources into oracle/_ref.

A space-frame tower under nonlinear Newmark (12 time steps, CHKPT 4) runs on
the reference's element routines under the restated driver for 10 steps,
writing a full-precision checkpoint at step 8.  Then:

  tower_ref_read.npz   what the reference's restartStep reads from that file
                       (main.c:3381-3386 reads its first line, tstep)
  tower_results8.txt   the file the reference's checkPoint writes from those
                       values: the checkpoint in the reference's own format
  tower_restart.npz    the path of the restarted run (RFLAG 1) from
                       tower_results8.txt, steps 9-12, on the reference's
                       element routines
  tower_model.txt, tower_model_restart.txt   the two model files

Run in the container that holds the reference:

    python tests/golden/make_checkpoint.py
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "cu-bens_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

OUT = os.path.join(HERE, "checkpoint")
TOWER = dict(n_elem=60, kind="tower", steps=12, amp=2e4, chkpt=4)
# the arrays ref_checkpoint / ref_restart take, in their argument order
ORDER = ("uc", "vc", "ac", "ss", "sm", "d", "f", "ef", "x", "c1", "c2", "c3", "defllen",
         "llength", "efFE", "xfr", "yldflag", "deffarea", "defslen")


def main():
    import ctypes as C

    import oracle as O
    from cubens import inputs
    if not O.ref_available():
        raise SystemExit("oracle/_ref is not built (make -C oracle with /root/reference present)")
    os.makedirs(OUT, exist_ok=True)
    tmp = tempfile.mkdtemp(prefix="cubens_ckpt_")
    base = inputs.space_frame(**TOWER).text
    restart = inputs.space_frame(**TOWER, rflag=1).text
    pb = os.path.join(OUT, "tower_model.txt")
    pr = os.path.join(OUT, "tower_model_restart.txt")
    open(pb, "w").write(base)
    open(pr, "w").write(restart)

    exact = os.path.join(tmp, "exact.txt")
    om = O.OracleModel(pb)
    r = om.run(10, backend="ref", ckpt_path=exact, ckpt_exact=True)
    assert r["nsteps"] == 10, r["msg"]

    O.RefModel(pb)                     # the reference's globals for this model
    sizes = O.ckpt_sizes(om)
    arrs = {k: np.zeros(max(sizes[k], 1), np.int32 if k == "yldflag" else np.float64) for k in ORDER}
    ptrs = [O.ptr(arrs[k], O.PI) if k == "yldflag" else O.ptr(arrs[k]) for k in ORDER]
    tstep = C.c_long()
    assert O.ref().ref_restart(exact.encode(), C.byref(tstep), om.lss, *ptrs) == 0
    assert tstep.value == 8
    np.savez_compressed(os.path.join(OUT, "tower_ref_read.npz"), tstep=tstep.value,
                        **{k: arrs[k][:sizes[k]] for k in ORDER})

    assert O.ref().ref_checkpoint(tmp.encode(), tstep.value, om.lss, *ptrs) == 0
    with open(os.path.join(tmp, "results8.txt")) as fh, open(os.path.join(OUT, "tower_results8.txt"), "w") as out:
        out.write(fh.read())

    ck = os.path.join(tmp, "restart_from.txt")
    with open(os.path.join(OUT, "tower_results8.txt")) as fh, open(ck, "w") as out:
        out.write(fh.read())
    rr = O.OracleModel(pr).run(13, backend="ref", ckpt_path=ck)
    assert rr["status"] == 0 and rr["nsteps"] == 4, rr["msg"]
    np.savez_compressed(os.path.join(OUT, "tower_restart.npz"), lpf=rr["lpf"], itecnt=rr["itecnt"],
                        d=rr["d"])
    for f in os.listdir(tmp):
        os.remove(os.path.join(tmp, f))
    os.rmdir(tmp)
    print("checkpoint fixtures written to", OUT)


if __name__ == "__main__":
    main()
