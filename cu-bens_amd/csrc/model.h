// Host-side model: CU-BENs model_def.txt input, DOF numbering, skyline profile,
// initial element geometry, and the node-block sparsity pattern with the element
// colouring used for atomic-free assembly on the device.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace cb {

struct Model {
    // control flags (main.c:340-427)
    int anaflag = 0, algflag = 0, slvflag = 0, optflag = 0, chkpt = 0, rflag = 0;
    long nj = 0, ntr = 0, nfr = 0, nsh = 0, neq = 0;
    // topology, 1-based as in the reference (model.c:42-1050)
    std::vector<long> minc, jcode, mcode, maxa, kht;
    std::vector<long> jnt;  // input joint -> joint number used (0-based; OPTFLAG 2 renumbers)
    std::vector<int> wrpres;
    long lss = 0;
    std::vector<double> x;  // NJ*3
    // element properties in the reference's flat layouts
    std::vector<double> emod, yield, carea, llength, dens_tr;
    std::vector<double> gmod, dens_fr, istrong, iweak, ipolar, iwarp, zstrong, zweak;
    std::vector<double> offset, auxpt, xfr;
    std::vector<int32_t> relmask;   // per frame member: its release flags as a mask (bit a: flag a)
    std::vector<int> osflag;
    std::vector<double> nu, thick, dens_sh, xlocal, farea, slength;
    std::vector<double> dens;        // the reference's one density array (see load_model)
    std::vector<double> c1, c2, c3;  // [NE_TR | NE_FR*3 | NE_SH*3]
    std::vector<double> q;           // NEQ reference load
    // dynamic input
    long ntstpsinpt = 0;
    double ttot = 0, dt = 0;
    std::vector<double> pinpt, um, vm, am;
    // driver parameters
    double lpfmax = 0, lpf0 = 0, dlpf0 = 0, dlpfmax = 0, dlpfmin = 0;
    int itemax = 0, submax = 0, solmin = 0, imagmax = 0, negmax = 0, iteopt = 0;
    double toldisp = 0, tolforc = 0, tolener = 0;
    double alpha = 0, psi_thresh = 0, dkimax = 0, dk = 0;
    long dkdof = 0;  // 0-based
    int numopt = 0;
    double spectrds = 1;

    // ---- device-facing pattern --------------------------------------------
    int bs = 6;                       // DOF slots per joint in the solver layout
    std::vector<int32_t> eq_of_slot;  // NJ*bs: 0-based equation or -1 (fixed)
    std::vector<int32_t> rowptr, colidx;  // joint-block CSR (both triangles)
    std::vector<int32_t> diagblk;         // block index of (j,j)
    // per element: block index of each ordered joint pair (a,b), a,b < nodes
    std::vector<int32_t> blk_tr;  // NE_TR*4
    std::vector<int32_t> blk_fr;  // NE_FR*4
    std::vector<int32_t> blk_sh;  // NE_SH*9
    // colour classes: elements of one class share no joint
    std::vector<int32_t> col_ptr_tr, col_el_tr, col_ptr_fr, col_el_fr, col_ptr_sh, col_el_sh;
    // row joint of every block, and per block the frame contributions that add
    // into it, in colour order: item = member * 4 + end_row * 2 + end_col
    std::vector<int32_t> blkrow, gat_ptr_fr, gat_item_fr;
    // per joint, the frame member ends at it in colour order: item = member * 2 + end
    std::vector<int32_t> jnt_ptr_fr, jnt_item_fr;
    long bandwidth = 0;  // max column height of the skyline
};

// Parses model_def.txt; returns false with a message on error.
bool load_model(const char* path, Model& m, std::string& err);
// Builds the joint-block pattern, the per-element block map and the colouring.
void build_pattern(Model& m);

}  // namespace cb
