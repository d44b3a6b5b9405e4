// mttkrp.hpp -- fused MTTKRP (tensor unfolding times on-the-fly Khatri-Rao product).
#pragma once
#include "common.hpp"

namespace tdx {

// Factor "panel" layout: rows grouped by 8, then column, then the 8 rows:
//   panel[((i >> 3) * Rp + r) * 8 + (i & 7)] = U[i + r*I]   (zero padded to Rp columns
//   and to a multiple of 64 rows)
// so one lane reads the 8 consecutive rows of one column with two 16-byte loads and
// 16 lanes of consecutive columns read 512 contiguous bytes.
inline long panel_rows(long I) { return ceil_div(I, 64) * 64; }
inline long panel_floats(long I, int Rp) { return panel_rows(I) * (long)Rp; }

struct MttkrpPlan {
    int order = 0;
    long dims[4] = {1, 1, 1, 1};
    int mode = 0;
    int dtype = TDX_F32;
    long rank = 0;
    int Rp = 0;          // rank padded to a multiple of 16
    bool fast = false;   // MFMA register-tile kernel usable
    bool lds = false;    // LDS-staged kernel (output mode > 0); ksteps = slow indices
    int rows_per_wg = 0; // output rows per workgroup tile
    int rowtiles = 0;
    long rows_pad = 0;   // rowtiles * rows_per_wg
    int nsplit = 0;      // K splits (workgroups per row tile)
    long ksteps = 0;
    long ksteps_per_split = 0;
    int lds_cps = 1;     // LDS kernel: i_0 chunks per split
    int lds_nss = 1;     // LDS kernel: slow-index ranges per chunk group
    bool split = false;  // BF16 storage, LDS kernel: Khatri-Rao block as hi + lo BF16 parts
    size_t partial_floats() const { return (size_t)nsplit * rows_pad * Rp; }
};

// chooses kernel variant and split; returns TDX_OK or an error.  split: the plan is for the
// hi + lo Khatri-Rao block (BF16 storage; the caller sets p.split when the LDS kernel is chosen)
int mttkrp_plan(MttkrpPlan &p, int order, const long *dims, int mode, int dtype, long rank,
                const void *X, int target_wgs = 0, bool split = false);

// Launches the split-K MTTKRP writing partial[split][row][Rp] (FP32).
// panels[m] are factor panels (panels[mode] unused); for FP8 storage panels[0] must be
// pre-scaled by the slice scales when mode != 0, and row_scale = scales when mode == 0.
int mttkrp_launch(const MttkrpPlan &p, const void *X, const float *const *panels,
                  const float *row_scale, float *partial, hipStream_t s, const int *skip = nullptr);

// the LDS-staged kernel (mttkrp_lds.hip)
bool mttkrp_lds_covers(int order, int mode, long rank, long In, int dtype);
int mttkrp_lds_launch(const MttkrpPlan &p, const void *X, const float *const *panels, const float *row_scale,
                      float *partial, hipStream_t s, const int *skip);

// the shared pass of an order-3 CP-ALS sweep (dtree.hip): T = X x_0 A^T with the mode-1
// partials in the same pass over X, then the mode-2 partials from T
struct DtreePlan {
    bool on = false;  // eligible (order 3, BF16/FP8 storage, I0 = 128 nch (BF16) / 256 nch (FP8)
                      // with nch in {2, 4, 8, 16}, I1 % 32 == 0, R <= 64, A operands fit in LDS)
    int st = 0, nrb = 0, nch = 0, Rp = 0, R = 0;
    long I0 = 0, I1 = 0, I2 = 0;
    int blocks_i1 = 0, nsplit1 = 0, i2per = 0;  // k_ttm_m1 grid: blocks_i1 x nsplit1
    size_t lds1 = 0;
    int nsplit2 = 0, tper = 0, nb2 = 0;         // k_m2 grid: nb2 x nsplit2, tper T tiles per split
    // T: BF16, I2 x (I1 / 32) tiles of 32 i1 x Rp, in the register order of the pass over X:
    // element (i1, r) of slab i2 at dt_t_index
    size_t t_elems() const { return (size_t)I1 * I2 * Rp; }
    size_t t_bytes() const { return 2 * t_elems(); }
    size_t partial_floats() const {
        const size_t a = (size_t)nsplit1 * I1 * Rp, b = (size_t)nsplit2 * I2 * Rp;
        return a > b ? a : b;
    }
};
// Element (i1, r) of slab i2 of T: tile (i2, i1 / 32) holds 32 Rp values; in it row block
// rb = r / 16 is 1 KB, lane l = r % 16 + 16 ((i1 % 16) / 4) of the pass over X holds 16 bytes
// of it, ordered c = (i1 % 32) / 16, then v = i1 % 4 (the MFMA output order of that pass).
__host__ __device__ inline long dt_t_index(long i1, long i2, int r, long I1, int Rp) {
    return (i2 * (I1 / 32) + i1 / 32) * 32L * Rp + (r / 16) * 512L + ((r % 16) + 16 * ((i1 % 16) / 4)) * 8L +
           ((i1 % 32) / 16) * 4 + i1 % 4;
}
// fills d (d.on = false when not enabled or when the shape or storage is not covered)
int dtree_plan(DtreePlan &d, int order, const long *dims, int dtype, long R, bool enabled);
// partial[split][i1][Rp] (nsplit1 splits, rows_pad I1) and T; C column-major I2 x R
int dtree_m1_launch(const DtreePlan &d, const void *X, const float *panelA, const float *C, uint16_t *T,
                    float *partial, hipStream_t s);
// partial[split][i2][Rp] (nsplit2 splits, rows_pad I2) from T and the mode-1 panel
int dtree_m2_launch(const DtreePlan &d, const uint16_t *T, const float *panelB, float *partial, hipStream_t s);
// the pass over T with the mode-2 inverse (G0 * G1)^{-1} -> Vinv in the same launch
// (16 < R <= 64)
int dtree_m2_vinv_launch(const DtreePlan &d, const uint16_t *T, const float *panelB, float *partial,
                         const double *G0, const double *G1, int R, double *Vinv, int *status, double *fb_ws,
                         hipStream_t s);

// panel packing (optionally scaling row i by scale[i])
int pack_panel(const float *U, long I, long R, int Rp, const float *scale, float *panel, hipStream_t s);

// sums the split partials into out (I_mode x rank, column-major FP32)
int mttkrp_reduce(const MttkrpPlan &p, const float *partial, float *out, hipStream_t s);

}  // namespace tdx
