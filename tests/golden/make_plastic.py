"""Writes tests/golden/plastic.npz: inelastic analyses (ANAFLAG 3) of the
restated drivers running on the reference's own element routines (compiled
from its sources into oracle/_ref), whose forces_fr carries the yield checks,
regula_falsi and unload of frame.c:1153-1270, 1397-1658 and whose stiff_fr /
stiff_tr carry stiffm_fr / stiffm_tr.

    tb_nr       three-bar truss, NR: the vertical bar capped at its squash load
    pc_nr       propped cantilever, NR: the clamped end yields (one
                regula-falsi restart) and the load goes on toward collapse
    pc_mnr      the same under modified NR (ALGFLAG 2)
    pc_msal     the same under MSAL through yielding, an unloading end and
                the collapse plateau, where the subdivisions run out
    pc_nm       the same under nonlinear Newmark, a ramped load that yields
    pc_nm_fail  a larger ramp: an unloading end, then the minimum time
                increment is reached without convergence
    sample_5b   the reference's sample input model_def_5b_frame.txt

Each case keeps the model text, the status, the path (lpf or time, itecnt, d),
the element forces and the driver events of every row.  Run in the
container that holds the reference:

    python tests/golden/make_plastic.py
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "cu-bens_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

MAX_STEPS = 250


def models() -> dict:
    from cubens import inputs
    pc = inputs.propped_cantilever().info["Pc"] / 1000.0
    with open(os.path.join(HERE, "model_def_5b_frame.txt")) as fh:
        s5b = fh.read()
    return {
        "tb_nr": inputs.three_bar_truss().text,
        "pc_nr": inputs.propped_cantilever(lpfmax=0.98 * pc, dlpf=0.05 * pc).text,
        "pc_mnr": inputs.propped_cantilever(lpfmax=0.98 * pc, dlpf=0.05 * pc, algflag=2).text,
        "pc_msal": inputs.propped_cantilever(analysis="msal", alpha=20.0, dk=-0.005).text,
        "pc_nm": inputs.propped_cantilever(analysis="newmark", load=-1.62e5, steps=40, dt=1e-3,
                                           tol=(1e-8, 1e-8, 1e-8)).text,
        "pc_nm_fail": inputs.propped_cantilever(analysis="newmark", load=-1.65e5, steps=40, dt=1e-3,
                                                tol=(1e-8, 1e-8, 1e-8)).text,
        "sample_5b": s5b,
    }


def main():
    import oracle as O
    if not O.ref_available():
        raise SystemExit("oracle/_ref is not built (make -C oracle with /root/reference present)")
    tmp = tempfile.mkdtemp(prefix="cubens_plastic_")
    out = {}
    for name, text in models().items():
        path = os.path.join(tmp, f"{name}.txt")
        with open(path, "w") as fh:
            fh.write(text)
        r = O.OracleModel(path).run(MAX_STEPS, backend="ref", keep_ef=True)
        out[f"{name}_model"] = np.array(text)
        out[f"{name}_status"] = np.array(r["status"])
        for k in ("lpf", "itecnt", "d", "ef", "events"):
            out[f"{name}_{k}"] = r[k]
        print(f"{name}: status {r['status']}, {r['nsteps']} rows, events {r['events'].sum(axis=0)}")
        os.remove(path)
    os.rmdir(tmp)
    np.savez_compressed(os.path.join(HERE, "plastic.npz"), **out)
    print("plastic.npz written")


if __name__ == "__main__":
    main()
