// Checkpoint file of the nonlinear Newmark driver: the state checkPoint writes
// and restartStep reads (misc.c:494-716), in the reference's layouts and line
// format, plus an optional full-precision extension after the reference's
// closing line (the working lengths and areas the mass routines overwrite and
// the time-increment multiplier of the next step), which the reference's
// reader never reaches.
#pragma once

#include <string>
#include <vector>

#include "model.h"

namespace cb {

struct Checkpoint {
    long tstep = 0;
    std::vector<double> uc, vc, ac;          // NEQ
    std::vector<double> ss, sm;              // lss: tangent K of the last iteration, lumped mass
    std::vector<double> d, f;                // NEQ
    std::vector<double> ef, x;               // NE_TR*2+NE_FR*14+NE_SH*18, NJ*3
    std::vector<double> c1, c2, c3;          // NE_TR+NE_FR*3+NE_SH*3
    std::vector<double> defllen, llength;    // NE_TR+NE_FR (llength: frame part in the file)
    std::vector<double> efFE, xfr;           // NE_FR*14, NE_FR*6
    std::vector<int> yldflag;                // NE_FR*2
    std::vector<double> deffarea, defslen;   // NE_SH, NE_SH*3
    bool has_ext = false;
    double ddt = 1;
    std::vector<double> llen_w, farea_w, slen_w;   // NE_TR+NE_FR, NE_SH, NE_SH*3

    void size_for(const Model& m);
};

// %e per value as the reference writes (exact = false), or %.17g and the
// extension.  Returns false when the file cannot be written.
bool write_checkpoint(const std::string& path, const Model& m, const Checkpoint& c, bool exact);
// 0 read, 1 cannot open, 2 malformed.  c is sized for m.
int read_checkpoint(const std::string& path, const Model& m, Checkpoint& c);

// The linear Newmark driver's checkpoint (solve.c:445-457, read at 257-282):
// the time step, NEQ lines "u,v,a", "0,0,0".
bool write_checkpoint_linear(const std::string& path, long k, const std::vector<double>& u,
                             const std::vector<double>& v, const std::vector<double>& a, bool exact);
int read_checkpoint_linear(const std::string& path, long neq, long* k, std::vector<double>& u,
                           std::vector<double>& v, std::vector<double>& a);

}  // namespace cb
