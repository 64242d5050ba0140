// eri_classes.hpp -- the 55 canonical angular-momentum classes (ss|ss)..(ff|ff).
//
// A shell pair class is (la, lb) with la >= lb, numbered pc = la(la+1)/2 + lb
// (ss=0, ps=1, pp=2, ds=3, dp=4, dd=5, fs=6, fp=7, fd=8, ff=9).  A quartet class
// pairs a bra class with a ket class of equal or lower number:
// qc = pb(pb+1)/2 + pk, 55 classes in all.
#pragma once

namespace hfx {

constexpr int MAX_L = 3;
constexpr int NPAIRCLASS = 10;
constexpr int NQCLASS = 55;

__host__ __device__ constexpr int pc_la(int pc) { return pc >= 6 ? 3 : pc >= 3 ? 2 : pc >= 1 ? 1 : 0; }
__host__ __device__ constexpr int pc_lb(int pc) { return pc - pc_la(pc) * (pc_la(pc) + 1) / 2; }
__host__ __device__ constexpr int pair_class(int la, int lb) { return la * (la + 1) / 2 + lb; }
__host__ __device__ constexpr int qc_bra(int qc) {
    int b = 0;
    while ((b + 1) * (b + 2) / 2 <= qc) ++b;
    return b;
}
__host__ __device__ constexpr int qc_ket(int qc) { return qc - qc_bra(qc) * (qc_bra(qc) + 1) / 2; }
__host__ __device__ constexpr int quartet_class(int pb, int pk) { return pb * (pb + 1) / 2 + pk; }

template <int QC>
struct QClass {
    static constexpr int LA = pc_la(qc_bra(QC)), LB = pc_lb(qc_bra(QC));
    static constexpr int LC = pc_la(qc_ket(QC)), LD = pc_lb(qc_ket(QC));
};

}  // namespace hfx
