// Internal (non-ABI) declarations shared by the host files of libctprj.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/ctprj.h"
#include "ctprj_device.h"

namespace ctprj {

struct BhcStore;  // bhc_internal.h

// A sum the symmetric adjoint's first wave adds on the side (*out = sum of in[0..n) in a
// fixed order), so that a reduction of the previous launches costs no launch of its own.
struct RedJob {
    const double* in = nullptr;
    long long n = 0;
    double* out = nullptr;
};

struct FwdItem {
    int slice, view, bin0, pad;
};
// partial-sum space of split symmetric work items: at most 256 MiB per plan
constexpr size_t kSplitMaxFloats = (size_t)64 << 20;
struct AdjItem {
    int slice, row0, col0, pad;
};

int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);

// Device tables of a plan.  The geometry tables depend on tune.flags, the work lists on
// tune.flags and the general kernels' shapes (fwd_block, adj_tile).  A tune change builds
// the parts it changes into a second Tables and swaps them in only when every step
// succeeded, so a failed change leaves the plan's tables as they were.
struct Tables {
    SliceDesc* d_desc = nullptr;
    float4* d_views = nullptr;
    Ray* d_rays = nullptr;
    int* d_symtab = nullptr;          // view-row table of the symmetric adjoint
    AdjItem* d_tr_items = nullptr;    // transpose tiles
    int n_tr_items = 0, n_sym = 0;
    std::vector<int> slice_sym;       // per slice: symmetry group (kGrp*), 0 = general kernels
    FwdItem* d_fwd_items = nullptr;   // general forward
    FwdItem* d_symf_items = nullptr;  // symmetric forward
    int n_fwd_items = 0, n_symf_items = 0;
    AdjItem* d_adj_items = nullptr;   // general adjoint
    AdjItem* d_syma_items = nullptr;  // symmetric adjoint
    int n_adj_items = 0, n_syma_items = 0;
    // symmetric work lists are grouped by symmetry group: [start, start + count) of
    // group g (1..3) in d_symf_items / d_syma_items
    int symf_start[4] = {0, 0, 0, 0}, symf_count[4] = {0, 0, 0, 0};
    int syma_start[4] = {0, 0, 0, 0}, syma_count[4] = {0, 0, 0, 0};

    Tables() = default;
    Tables(const Tables&) = delete;
    Tables& operator=(const Tables&) = delete;
    ~Tables() { release(); }
    void release();
    void swap_geometry(Tables& o);
    void swap_fwd(Tables& o);
    void swap_adj(Tables& o);
};

struct Plan : Tables {
    int device = 0;
    int nslices = 0;
    std::vector<ctprj_geom> geoms;
    std::vector<long long> img_offsets, sino_offsets;
    long long img_elems = 0, sino_elems = 0;
    int max_W = 0;
    ctprj_tune tune{64, 16, 1, 8, 8, 1, 1};

    int2* d_fbp_items = nullptr;
    float* d_zero_row = nullptr;      // max_W zeros: reads of views beyond np
    float* d_fbp_kern = nullptr;
    long long* d_fbp_kern_off = nullptr;
    float* d_fbp_prew = nullptr;
    long long* d_fbp_prew_off = nullptr;
    int n_fbp_items = 0;

    float* d_imgT = nullptr;     // transposed image (forward, steep views / symmetric images)
    float* d_sino_ws = nullptr;  // sinogram workspace (FBP filter output, BHC ray values)
    double* d_red = nullptr;     // reduction partials (beam-hardening)
    double* d_red_out = nullptr; // reduced sums
    long long red_cap = 0;       // doubles held by d_red
    int bhc_basis = CTPRJ_BASIS_B1, bhc_noise = CTPRJ_NOISE_POISSON;  // spectrum model
    void* d_sim = nullptr;       // measurement simulator workspace
    size_t sim_cap = 0;          // bytes held by d_sim
    float* d_ie_phi = nullptr;   // I-epsilon step: basis matrix of its projection
    size_t ie_phi_cap = 0;       // bytes held by d_ie_phi
    float* d_split = nullptr;    // partial sums of split symmetric work items
    size_t split_cap = 0;        // floats held by d_split

    ~Plan();
    // geometry tables for tune flags `flags` into G
    int build_geometry(int flags, Tables& G) const;
    // work lists for tune t on the geometry of G, into G
    int build_fwd_items(const ctprj_tune& t, Tables& G) const;
    int build_adj_items(const ctprj_tune& t, Tables& G) const;
    // partial-sum floats the split symmetric launches of tune t need on tables G
    size_t split_need(const ctprj_tune& t, const Tables& fwd, const Tables& adj) const;
    int launch_forward(const float* d_img, float* d_sino, hipStream_t st, bool count);
    // Number of NLL partials of launch_forward_bhc (the symmetric forward items), or -1
    // when the plan has slices on the general kernels (no fused store there).
    long long bhc_fused_parts() const;
    // Forward projection with the alpha pass fused into the store (bhc->partials gets
    // bhc_fused_parts() values from offset 0; bhc->part_off is set here).
    int launch_forward_bhc(const float* d_img, float* d_w, BhcStore bhc, hipStream_t st);
    // job: a sum done by the first symmetric launch (none if job.out is NULL; the plan
    // must then have a symmetric adjoint launch, see has_sym_adjoint)
    int launch_adjoint(const float* d_sino, float* d_img, hipStream_t st, bool fbp,
                       const RedJob& job = RedJob{});
    bool has_sym_adjoint() const { return n_syma_items > 0; }
    int launch_fbp(const float* d_sino, float* d_img, hipStream_t st);
};

// Kernel launchers (ctprj_kernels.hip).
void launch_build_rays(const SliceDesc* desc, int nslices, const float4* views,
                       const float4* bins, const long long* bin_off, const long long* ray_prefix,
                       Ray* rays, long long nray, hipStream_t st);
void launch_transpose(int nitems, const SliceDesc* desc, const AdjItem* items, const float* in,
                      float* out, hipStream_t st);
void launch_forward_kernel(bool count, int block, int nitems, const SliceDesc* desc,
                           const FwdItem* items, const float4* views, const Ray* rays,
                           const float* img, const float* imgT, float* sino, hipStream_t st);
void launch_adjoint_kernel(int tile, bool fbp, int nitems, const SliceDesc* desc,
                           const AdjItem* items, const float4* views, const Ray* rays,
                           const float* sino, float* img, hipStream_t st);
void launch_fbp_filter(int nitems, size_t shmem, const SliceDesc* desc, const int2* items,
                       const float* kern, const long long* kern_off, const float* prew,
                       const long long* prew_off, const float* sino, float* out,
                       hipStream_t st);
void launch_scale(float* p, long long n, float s, hipStream_t st);
// symmetric kernels (ctprj_sym.hip), one launch per symmetry group; nsplit > 1 work-groups
// per item leave partial sums in part (nitems * nsplit * 512 floats), added by a second
// launch.  bhc != NULL: the beam-hardening alpha pass fused into the store (sino receives
// w; one NLL partial per item at bhc->partials[bhc->part_off + item])
void launch_forward_sym(int grp, int seg, int nsplit, float* part, int nitems,
                        const SliceDesc* desc, const FwdItem* items, const float4* views,
                        const Ray* rays, const float* img, const float* imgT, float* sino,
                        const BhcStore* bhc, hipStream_t st);
// The same sum as a symmetric adjoint's side job, in a launch of its own.
void launch_wave_sum(const RedJob& job, hipStream_t st);
void launch_adjoint_sym(bool fbp, int grp, int seg, int win, bool stage, int nsplit, float* part,
                        int nitems, const SliceDesc* desc, const AdjItem* items,
                        const float4* views, const Ray* rays, const float* sino,
                        const float* zero_row, const int* symtab, float* img, const RedJob& job,
                        hipStream_t st);

// Images held by a wavelet context (rows * cols * batch).
long long wavelet_elems(const ctprj_wavelet* w);
// Images held by a TV context (rows * cols * batch).
long long tv_elems(const ctprj_tv* w);
// Prior penalties in two steps: queue the per-group partial sums of ||W x||_1 or TV(x)
// into d_part (*_parts doubles), then, once they are in host memory, add them in the
// order ctprj_wavelet_l1 / ctprj_tv_value use.
long long wavelet_l1_parts(const ctprj_wavelet* w);
int wavelet_l1_queue(ctprj_wavelet* w, const float* d_x, double* d_part, hipStream_t st);
double wavelet_l1_total(const ctprj_wavelet* w, const double* part);
long long tv_value_parts(const ctprj_tv* w);
int tv_value_queue(ctprj_tv* w, const float* d_x, double* d_part, hipStream_t st);
double tv_value_total(const ctprj_tv* w, const double* part);

}  // namespace ctprj
