// The elementwise vector algebra of the nonlinear Newmark driver, entry i of
// an equation-order vector.  One definition serves the single-model launches
// and the batched ones, so a model gets the same arithmetic in both.
#pragma once

#include <hip/hip_runtime.h>

#include "dev.h"

namespace cb {

__device__ inline void nm_apply(NewmarkVec op, const NewmarkArgs& A, long i) {
    switch (op) {
    case NM_PREDICT0: {
        // main.c:3553-3560
        const double q = A.P[i * A.ld + A.k] * A.sub_dt * A.lpf;
        A.qtot[i] = q;
        A.r[i] = q - A.ft[i];
        break;
    }
    case NM_PREDICT: {
        // main.c:3561-3569
        const double p0 = A.P[i * A.ld + A.k - 1], p1 = A.P[i * A.ld + A.k];
        const double q = (p0 + (p1 - p0) * A.sub_dt) * A.lpf;
        A.qtot[i] = q;
        A.r[i] = (q - A.ft[i]) + A.af / (1 - A.af) * (p0 - A.ft[i]);
        break;
    }
    case NM_CORRECT: {
        // main.c:3572-3576
        double r = A.ft[i] - A.qtot[i];
        A.r[i] = r + A.sm[i] * A.ac[i] -
                 A.sm[i] * ((A.a2 * A.vc[i] + A.a3 * A.ac[i]) * (1 - A.am) - A.ac[i] * A.am) / (1 - A.af);
        break;
    }
    case NM_REFF: {
        // solve.c:482-484
        A.out[i] = A.r[i] + A.sm[i] * ((1 - A.am) * (A.vc[i] * A.a2 + A.ac[i] * A.a3) - A.am * A.ac[i]) /
                                (1 - A.af);
        break;
    }
    case NM_UPDATE: {
        // main.c:3631-3653 (dd sign flip on corrector iterations, kinematics)
        double ddi = A.dd[i];
        if (A.flip) ddi = -ddi;
        A.dd[i] = ddi;
        const double u = A.dt[i] + ddi;
        A.dt[i] = u;
        A.fip[i] = A.ft[i];
        A.ft[i] = 0.0;
        const double a = (u - A.um[i]) * A.a0 - A.a2 * A.vm[i] - A.a3 * A.amv[i];
        A.ac[i] = a;
        A.vc[i] = A.vm[i] + A.a6 * A.amv[i] + A.a7 * a;
        break;
    }
    case NM_DYN: {
        // main.c:3684-3688
        A.out[i] = A.qtot[i] - A.sm[i] * A.ac[i];
        break;
    }
    }
}

}  // namespace cb
