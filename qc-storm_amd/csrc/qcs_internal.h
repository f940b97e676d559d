// qcs_internal.h -- shared declarations of libqcstorm (not part of the public C-ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include "../../include/qcstorm.h"

namespace qcs {

// Thread-local error text; returns `code`.
int fail(int code, const char *fmt, ...);
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define QCS_HIP(call)                                                        \
    do {                                                                     \
        hipError_t _e = (call);                                              \
        if (_e != hipSuccess) return qcs::hip_fail(_e, #call, __FILE__, __LINE__); \
    } while (0)

// Scalar fit configuration passed to kernels by value (CoreFittingPara, bfgs_base.h:113-141).
struct FitConst {
    float Offset, KAdc, QE, ReadNoise_e, PixelSize;
    int MultiEmitterFitEn, WLEEn;
    float ZDepthCorrFactor;
    float pG[5];   // p4..p0 for sigma_x >= sigma_y
    float pL[5];   // p4..p0 for sigma_x <  sigma_y
};
FitConst make_fit_const(const qcs_params &p);

// Grow-only device buffer.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t b);
    void release();
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// Pinned host buffer.
struct HostBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t b);
    void release();
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

int validate_params(const qcs_params &p);

// stage timing (api.cpp): a recorded interval per launch group, resolved on demand
enum { TM_EXTRACT = 0, TM_FIT1E, TM_FITME, TM_CRLB, TM_COMPACT, TM_RENDER, TM_CFIT, TM_NN, TM_DRIFTCC, TM_COUNT };
struct Timer {
    int on = 0;
    struct Pending {
        int cat;
        hipEvent_t a, b;
        long long items;
    };
    Pending *pend = nullptr;
    int npend = 0, cap = 0;
    double ms[TM_COUNT] = {0};
    long long launches[TM_COUNT] = {0}, items[TM_COUNT] = {0};
};
// begin/end an interval on stream s (no-ops when timing is off); end returns the status
int timer_begin(Timer &t, hipStream_t s, hipEvent_t *ev);
int timer_end(Timer &t, int cat, hipStream_t s, hipEvent_t ev, long long items);

// fit stage internals (fit.hip).  n is the grid bound; when d_n is set the kernels process
// *d_n items (a count that lives on the device).
int fit_launch_1e(const qcs_params &p, const uint16_t *d_rois, const float *d_wle, const int *d_idx, int n,
                  const int *d_n, float *d_loc, int *d_flag, hipStream_t s);
int fit_launch_me(const qcs_params &p, int model, const uint16_t *d_rois, const float *d_wle, const int *d_list,
                  int n_max, const int *d_n, float *d_loc, float *d_second, int *d_fs, int *d_ff, hipStream_t s);
int crlb_launch(const qcs_params &p, float *d_loc, int n, const int *d_n, hipStream_t s);

// Scratch of the multi-emitter cascade (fit.hip), carved from one block of
// cascade_scratch_bytes(n) bytes: counters (16 ints; [7] = rows written once the launches end),
// the 1E / 2E / 3E lists, flags and scan offsets, and the second-emitter rows of each stage.
struct CascadeBufs {
    int *cnt, *list1, *list2, *list3, *flag;
    float *rows;
};
size_t cascade_scratch_bytes(int n);
void cascade_carve(void *base, int n, CascadeBufs &b);
// Launch the whole cascade (fits, list building, appends, CRLB) on stream s without waiting;
// stage intervals go to *tm when it is given and enabled.
int fit_cascade_launch(const qcs_params &p, Timer *tm, const uint16_t *d_rois, const float *d_wle, int n, float *d_loc,
                       int loc_cap, const CascadeBufs &b, hipStream_t s);

// file-scope scratch of the allocation-free entry points (noalloc.hip): resolves the calling
// device's copy once per device (the address that host code passes to kernels)
int noalloc_scratch_init(int device);

// rendering helpers (post.hip): splat rows into the SR image(s); max of (int) image value into
// *d_out (which must start at 0)
int render_launch(qcs_ctx *ctx, const float *d_loc, int n, float *d_img, float *d_zimg, bool fresh, hipStream_t s);
int max_disp_launch(const float *d_img, size_t npx, int *d_out, hipStream_t s);

// device utilities (extract.hip): exclusive scan of n ints (total at tile_sums[ntiles]),
// constant fill, stable compaction out[base + k] = src ? src[j] : j for flagged j.
int scan_tile_count(size_t n);
int scan_exclusive(const int *in, int n, int *out, int *tile_sums, hipStream_t s);
void fill_async(int *p, size_t n, int v, hipStream_t s);
int compact_scatter(const int *flags, const int *offs, int n, const int *src, const int *base_dev, int base_const,
                    int *out, hipStream_t s);

// stable LSD radix sort of (key, int) pairs by the low `bits` key bits (sort.hip): k0/v0 hold the
// input, k1/v1 the same sizes; *kres/*vres receive whichever pair holds the result.  tmp holds
// radix_sort_tmp_ints(n) ints.  Inclusive prefix maximum of n ints (tmp: max_scan_tmp_ints(n)).
size_t radix_sort_tmp_ints(int n);
int radix_sort_pairs_u32(unsigned *k0, int *v0, unsigned *k1, int *v1, int n, int bits, int *tmp, hipStream_t s,
                         unsigned **kres, int **vres);
int radix_sort_pairs_u64(unsigned long long *k0, int *v0, unsigned long long *k1, int *v1, int n, int bits, int *tmp,
                         hipStream_t s, unsigned long long **kres, int **vres);
size_t max_scan_tmp_ints(int n);
int scan_inclusive_max(const int *in, int n, int *out, int *tmp, hipStream_t s);

}  // namespace qcs

struct qcs_ctx {
    qcs_params p;
    int device = 0;
    qcs::Timer timer;
    qcs_fit_stats stats{};
    // extraction scratch
    qcs::DevBuf ex_bkg, ex_tmp, ex_det, ex_mark, ex_cnt, ex_xy, ex_var, ex_frames;
    // fit scratch
    qcs::DevBuf fit_idx1, fit_idx2, fit_idx3, fit_flag, fit_cnt, fit_rows;   // fit_flag: the cascade's carve
    // converged fitting scratch
    qcs::DevBuf cf_theta, cf_ll, cf_list, cf_rows, cf_queue;
    // mixed-size extraction (7 x 7 cuts, WLE, size flags, set lists) and the frames path's sets
    qcs::DevBuf mx_rois, mx_list, mx_sets, mx_rows;
    // rendering scratch (tile bins), statistics scratch
    qcs::DevBuf rn_buf, st_buf;
    qcs::DevBuf dr_img, dr_cc, dr_shift, dr_frames;
    qcs::DevBuf dp_hist;                   // display image histogram   // drift correction: two SR images, FP64 window, per-frame shifts
    qcs_tuning tune = {{100, 40, 40}, 1e-4f, 0.8f, 1, 0, 1e-3f, 0, 0};
    // host staging for the host-pointer API
    qcs::DevBuf h_rois_dev, h_wle_dev, h_loc_dev;
    qcs::HostBuf pin_cnt;
    float taps[5];
};
