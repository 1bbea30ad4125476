// phf_tree.hpp -- uniform Morton quad-tree build (device).
#pragma once

#include "phf_ops.hpp"

namespace phf {

// Square domain [x0, x0+D) x [y0, y0+D) split into 2^L x 2^L leaf boxes.
struct Domain {
    double x0, y0;  // lower-left corner (bbox minimum)
    double D;       // side, bbox extent rounded up to 5 significant bits
    double s;       // leaf box side D / 2^L (exact)
    double inv_s;   // 2^L / D
    int L;
    // sort keys carry 2 * fine more bits below the leaf index, the Morton
    // order of a 2^fine x 2^fine grid inside each leaf (leaf = key >> 2 fine)
    int fine = 0;
    // the same corner on the device ({xmin, xmax, ymin, ymax} from
    // device_bbox): kernels read the origin there, so a captured launch
    // sequence stays valid when the points move within the same side
    const double4 *org = nullptr;
    // optional: the partial boxes of device_bbox_parts, reduced by the
    // first tree build that gets them, which then stores the box at *org
    const double4 *parts = nullptr;
    int nparts = 0;
    // with parts: also a host-coherent pinned copy of the completed box
    double4 *org_host = nullptr;
};

// side > 0: the domain side to use (at least the extent, from
// domain_side); otherwise the bounding-box side.
Domain make_domain(double xmin, double xmax, double ymin, double ymax, int L,
                   double side = 0.0);

// Exact bounding box of (x1,y1)[n1] U (x2,y2)[n2] into d_out = {xmin, xmax,
// ymin, ymax} (device memory).
int device_bbox(Lane *L, const double *x1, const double *y1, int64_t n1, const double *x2,
                const double *y2, int64_t n2, double4 *d_out);
// its first kernel only: *np partial boxes at *parts (arena)
int device_bbox_parts(Lane *L, const double *x1, const double *y1, int64_t n1, const double *x2,
                      const double *y2, int64_t n2, double4 **parts, int *np);

int device_exclusive_scan(Lane *L, const uint32_t *in, int64_t n, uint32_t *out);
// variants with caller-provided workspaces (no allocation: capturable)
int64_t scan_sums_size(int64_t n);
// dn (device, nullable): scan only the first *dn <= n entries
int device_exclusive_scan_ws(Lane *L, const uint32_t *in, int64_t n, uint32_t *out,
                             uint32_t *sums, const int64_t *dn = nullptr);
int64_t radix_ws_size(int64_t n);  // entries of hist and of offs
int device_radix_sort_ws(Lane *L, uint32_t *keys, int32_t *vals, uint32_t *ktmp, int32_t *vtmp,
                         int64_t n, int key_bits, uint32_t *hist, uint32_t *offs,
                         uint32_t **kres, int32_t **vres);

// Stable LSD radix sort of (keys, vals) on the low key_bits bits; ktmp/vtmp
// are ping-pong buffers.  *kres/*vres point at whichever buffer holds the
// result.
int device_radix_sort(Lane *L, uint32_t *keys, int32_t *vals, uint32_t *ktmp, int32_t *vtmp,
                      int64_t n, int key_bits, uint32_t **kres, int32_t **vres);
// The same with the key width known only on the device: key_bits passes'
// worth of launches, the passes at or above *dbits copying their tile
// unchanged (the result lands in the same buffer either way).
int device_radix_sort_dbits(Lane *L, uint32_t *keys, int32_t *vals, uint32_t *ktmp, int32_t *vtmp,
                            int64_t n, int key_bits, const int32_t *dbits, uint32_t **kres,
                            int32_t **vres);

struct SortedPoints {
    int64_t n = 0;
    uint32_t *keys = nullptr;     // sorted Morton keys
    int32_t *perm = nullptr;      // sorted position -> input index
    int32_t *start = nullptr;     // leaf offsets, 4^L + 1 entries
    uint32_t *keys_in = nullptr;  // keys in input order (if requested)
    bool want_input_keys = false;
    // build_sorted_payload: sorted float32 records (x - cx, y - cy, w,
    // input index bits) relative to the leaf centres, perm unset
    float4 *pay = nullptr;
    // leaf ranges [split[k], split[k+1]) of the 8 XCDs holding equal
    // numbers of points (9 entries; if requested)
    int32_t *split = nullptr;
    bool want_split = false;
    int fine = 0;  // keys hold 2 * fine bits below the leaf index
};

// start[k] = first position of sorted keys with (key >> 2 fine) >= k,
// k = 0..nkeys; split (nullable): the XCD leaf ranges as SortedPoints::split.
int leaf_starts_device(Lane *L, const uint32_t *skeys, int64_t n, int64_t nkeys, int32_t *start,
                       int32_t *split = nullptr, int fine = 0);

int build_sorted_points(Lane *L, const Domain &dom, const double *x, const double *y, int64_t n,
                        SortedPoints *out);
// The same order with the points' float32 records sorted along (no
// permutation gather afterwards).
int build_sorted_payload(Lane *L, const Domain &dom, const double *x, const double *y,
                         const double *w, int64_t n, SortedPoints *out);

}  // namespace phf
