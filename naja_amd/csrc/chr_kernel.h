// Internal interface between the C ABI (capi.cpp) and the CHR step kernel.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace naja {

// One conditioned variant, packed by chr_pack_kernel: rows [A; E_v] in a
// column-major block with lda = 64 W R (the launch shape's rows); padding
// rows hold A = 0 and b = +inf, so their slack stays +inf and their chord
// ratio (inf / 0, or inf * 0 when the register kernel's ranged loads read the
// padding as 0) never bounds the chord of a bounded polytope.
struct ChrVariant {
    int m;               // real packed rows (m0 + k; fewer with paired rows)
    int lda;             // padded rows (64 W R)
    int64_t a_off;       // float offset of the packed block
    int64_t b_off;       // offset of the packed FP64 right-hand side
    int64_t y0_col;      // first column of this variant's chains in y0 / y_final
    int64_t sample_off;  // float offset of this variant's sample block
};

// Work item: up to G chains of one variant that share the move stream.
struct ChrItem {
    int variant;
    int c0;  // first chain (a multiple of G)
};

struct ChrArgs {
    const float* A;     // packed variant blocks
    const float* Ainv;  // RN(1 / A) of the packed blocks (coordinate-move chords)
    const double* b;  // packed right-hand sides
    const double* bn;  // paired-row layout: the partner rows' right-hand sides (-a_r y <= bn_r); else null
    int d;
    const ChrVariant* variants;
    const ChrItem* items;
    const float* y0;
    float* samples;
    float* y_final;
    int nchains;
    int thin;
    int64_t nsteps;
    uint32_t seed_lo;
    uint32_t seed_hi;
    uint64_t thr_k;   // k-sparse move: w0 < thr_k
    uint64_t thr_kp;  // pair move: w0 < thr_kp (and not k-sparse)
    int has_pair;
    int pair_mode;
    const int* pair_i;
    const int* pair_j;
    const float* pair_c;
    const float* pair_s;
    int n_pairs;
    int ks_k;
    float ks_mag;
    int resync;
    int32_t* move_log;
    int64_t nlog;
    int64_t t0;  // global index of the first step (Philox counter); a multiple of thin
    // Scaled pairs (paired kernels): the packs hold each pair's rows divided
    // by a power of two S >= 2 (b + bn) and its reciprocals multiplied by S,
    // so slacks lie in [0, 1/2] and a clamp to [0, 1] is the FMA's own output
    // clamp; padding reciprocals are +inf.
    int scaled;
};

// FP64 resync window actually used: R > 0 as given; 0 = automatic
// max(1024, 16 d) steps (FP32 slack needs a bounded drift window, where the
// reference's FP64 slack does not, chr.cu:386); R < 0 = never after t = 0.
inline int chr_effective_resync(int requested, int d)
{
    if (requested > 0) return requested;
    if (requested < 0) return 0;
    return d * 16 > 1024 ? d * 16 : 1024;
}

// Launch shape: W wavefronts per workgroup, G chains per group, R row slots
// per lane; every packed block has exactly 64 W R rows.
struct ChrShape {
    int W, G, R;
    int rows() const { return 64 * W * R; }
    // The large-polytope kernel (16 wavefronts, more than 8 rows per lane)
    // forms reciprocals itself and reads no packed reciprocal column.
    bool big() const { return W == 16 && R > 8; }
    // Dynamic LDS: the G x d chain states.
    size_t lds_bytes(int d) const { return sizeof(float) * (size_t)G * d; }
};

using ChrKernelFn = void (*)(ChrArgs);
struct ChrKernelEntry {
    int W, G, R;
    ChrKernelFn fn;
    // Same shape specialised for launches whose every move is a coordinate
    // move (no pair or k-sparse probability); null when not compiled.
    ChrKernelFn fn_coord = nullptr;
    // One-wavefront shapes with the chain states in VGPRs (d <= 128), general
    // and coordinate-only; null when not compiled.
    ChrKernelFn fn_x2 = nullptr;
    ChrKernelFn fn_coord_x2 = nullptr;
    // The same four for the paired-row layout (a packed row and its negation
    // share one column entry and keep two slacks); null when not compiled.
    ChrKernelFn fn_p = nullptr;
    ChrKernelFn fn_coord_p = nullptr;
    ChrKernelFn fn_x2_p = nullptr;
    ChrKernelFn fn_coord_x2_p = nullptr;
    // Coordinate-only, VGPR states, four chains of one variant per workgroup
    // with the packed A in LDS (unpaired and paired layouts); null when not
    // compiled.
    ChrKernelFn fn_coord_x2_l4 = nullptr;
    ChrKernelFn fn_coord_x2_p_l4 = nullptr;
    // The paired kernels on scaled pairs (ChrArgs::scaled): twins of
    // fn_coord_x2_p_l4, fn_p, fn_coord_p, fn_x2_p and fn_coord_x2_p (the
    // tables initialise the members in this order).
    ChrKernelFn fn_coord_x2_ps_l4 = nullptr;
    ChrKernelFn fn_ps = nullptr;
    ChrKernelFn fn_coord_ps = nullptr;
    ChrKernelFn fn_x2_ps = nullptr;
    ChrKernelFn fn_coord_x2_ps = nullptr;
};
// An entry holding only scaled-pair kernels (the tables that compile them).
constexpr ChrKernelEntry chr_scaled_entry(int W, int R, ChrKernelFn ps, ChrKernelFn coord_ps = nullptr,
                                          ChrKernelFn x2_ps = nullptr, ChrKernelFn coord_x2_ps = nullptr)
{
    ChrKernelEntry e{W, 1, R, nullptr};
    e.fn_ps = ps;
    e.fn_coord_ps = coord_ps;
    e.fn_x2_ps = x2_ps;
    e.fn_coord_x2_ps = coord_x2_ps;
    return e;
}
// Chains per workgroup of the LDS-column kernels.
constexpr int kChrLdsChains = 4;
// Largest d of the register-resident-state kernels (two VGPRs per lane).
constexpr int kChrRegStateMaxD = 128;
struct ChrKernelTable {
    const ChrKernelEntry* e;
    int n;
};
// Tables of compiled shapes, one per chr_inst_*.hip translation unit.
ChrKernelTable chr_table_w1();
ChrKernelTable chr_table_w4a();
ChrKernelTable chr_table_w4b();
ChrKernelTable chr_table_w8();
ChrKernelTable chr_table_w16();
ChrKernelTable chr_table_g4();
ChrKernelTable chr_table_big();  // W = 16, R = 10..22: the large-polytope kernel
ChrKernelTable chr_table_w1p();  // one wavefront, paired rows
ChrKernelTable chr_table_w1l();  // one wavefront, four chains per workgroup, A columns in LDS
ChrKernelTable chr_table_w4p();  // four wavefronts, paired rows
ChrKernelTable chr_table_w8p();  // eight wavefronts, paired rows
ChrKernelTable chr_table_w1s();  // one wavefront, paired rows on scaled pairs
ChrKernelTable chr_table_w4s();  // four wavefronts, paired rows on scaled pairs
ChrKernelTable chr_table_w8s();  // eight wavefronts, paired rows on scaled pairs

// Smallest compiled R with 64 W R >= rows for W wavefronts and G chains
// per group (among the paired-row layout's kernels when paired); -1 if none.
int chr_smallest_r(int W, int G, int rows, bool paired = false);

// Shape for a batch whose largest variant has m_max packed rows (tpb > 0 is
// the reference's --tpb, taken as a preference). Paired: among the
// paired-row kernels, and W = 0 when none holds the rows.
ChrShape chr_choose_shape(int m_max, int G, int tpb, int n_items, bool paired = false);
// nc = kChrLdsChains launches the LDS-column kernel (items in groups of nc
// chains of one variant, lds_bytes of dynamic LDS); nc = 1 the others.
void chr_launch(const ChrArgs& args, int n_items, const ChrShape& sh, hipStream_t stream, int nc = 1,
                size_t lds_bytes = 0);
// Whether (and with how much LDS) a launch of shape sh can use the
// LDS-column kernel: one wavefront, one chain per group, coordinate moves
// only, d <= kChrRegStateMaxD, nchains a multiple of kChrLdsChains, and a
// variant's d columns of 64 W R packed rows within the LDS.
size_t chr_lds_columns_bytes(const ChrShape& sh, int d);
bool chr_has_lds_kernel(const ChrShape& sh, bool paired);
// Whether the kernel a launch of shape sh would run (coordinate moves only or
// not, chain states in VGPRs or not, nc chains per workgroup) has a
// scaled-pair twin.
bool chr_has_scaled_kernel(const ChrShape& sh, bool coord_only, bool reg_state, int nc);
// Launch timing for measurement: while enabled, chr_launch records HIP events
// around every step-kernel launch; collect returns (work items, ms) of up to
// max recorded launches, oldest first, and drops the record.
void chr_timing_enable(bool on);
int chr_timing_collect(int* items, float* ms, int max);

// Packs A (m0 x d, lda m0) + per-variant extra blocks into padded blocks;
// with bn_pack, in the paired-row layout (paired_dev: the flags of
// chr_pair_check).
// scaled: every block paired, the packs in the scaled-pair form of
// ChrArgs::scaled.
void chr_pack(const float* A, const double* b, int m0, int d, const float* A_extra, const double* b_extra,
              const ChrVariant* variants_dev, const int64_t* extra_off_dev, const int* extra_k_dev, int n_variants,
              const int* paired_dev, float* A_pack, float* Ainv_pack, double* b_pack, double* bn_pack, int lda_max,
              hipStream_t stream, bool scaled = false);

// Rounding transform: A' = A P_v, b' = b - A c_v per variant (see
// chr_kernel.hip); bn_pack / bn_out for the paired-row layout (else null).
void chr_transform(const float* A_pack, const double* b_pack, const double* bn_pack, const ChrVariant* variants_dev,
                   int n_variants, int d, int lda, const float* P, const float* c, float* A_out, float* Ainv_out,
                   double* b_out, double* bn_out, hipStream_t stream, bool scaled = false);

// Row pairing: flags_dev[0] (base block) and flags_dev[1 + v] (extra block of
// variant v), set to 1 by the caller, are cleared unless the block's second
// half is the negation of its first.
void chr_pair_check(const float* A, int m0, int d, const float* A_extra, const int64_t* extra_off_dev,
                    const int* extra_k_dev, int n_variants, int* flags_dev, hipStream_t stream);

}  // namespace naja
