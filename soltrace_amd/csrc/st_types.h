// Plain-old-data tables shared by the host scene compiler and the HIP
// kernels.  All device tables are flat arrays of these structs in HBM; the
// trace kernels stage the per-stage headers and small element tables in LDS.
#pragma once
#include <cstdint>

namespace st {

enum : int32_t { kMaxStages = 16 };

// Surface kinds after parameter translation (coretrace/input.cpp:288-366,
// surface type codes of coretrace/intersect.cpp:182-209).
enum SurfKind : int32_t {
    SURF_SPHERE = 1,      // 's'  closed form (type 1 or 7)
    SURF_PARABOLA = 2,    // 'p'  closed form (type 1 or 7)
    SURF_CONIC = 3,       // 'o'  closed form (type 1)
    SURF_CYLINDER = 4,    // 't'  closed form (type 2)
    SURF_TORUS = 5,       // 'd'  quartic     (type 10)
    SURF_FLAT = 6,        // 'f'  Newton      (type 3)
    SURF_ZERNIKE = 7,     // 'm'  Newton      (type 6)
    SURF_SPLINE = 8,      // 'i'  Newton      (type 9)
    SURF_GENERAL = 9,     // 'g'/'c'/'o'(type 7) Newton on the conic (type 1/7)
    SURF_POLY = 10,       // 'r'  Newton      (type 8)
    SURF_VSHOT = 11,      // 'v'  Newton      (type 5): monomial z, measured slopes
    SURF_FE = 12,         // 'e'  Newton      (type 4): kriging over finite-element nodes
};

enum ApKind : int32_t {
    AP_CIRCLE = 1, AP_HEX = 2, AP_TRI = 3, AP_RECT = 4, AP_ANNULUS = 5, AP_LINE = 6, AP_ITRI = 7, AP_QUAD = 8,
};

// One optical face (coretrace/types.h:85-109).  Tables live in the aux
// buffer as (angle, value) pairs.
struct DevFace {
    int32_t dist;           // 'g','p','d','f' as ASCII
    int32_t refl_tab, refl_n, trans_tab, trans_n;
    int32_t pad0;
    double reflectivity, transmissivity;
    double slope, spec;     // mrad
    double n_in, n_out;     // RefractiveIndex[0] and [2]
};

// One element.  Frame data is always double; surface and aperture
// parameters are double here and converted per precision in the kernels.
struct DevElem {
    double org[3];          // origin in stage frame
    double R[9];            // stage -> element rotation (row major)
    double p[8];            // aperture parameters A..H
    double zap;             // aperture-plane z (coretrace/apertureplane.cpp)
    double kappa, vcx, vcy;
    double ann_r, cross_r, curv_rev, cone_tan;  // cone_tan = tan(half angle)
    double bsphere_c[3];    // bounding sphere of accepted hits, element frame
    double bsphere_r;
    int32_t ap, surf, stype, interaction;
    int32_t face_front, face_back;  // indices into the face table (-1: none)
    int32_t aux, aux_n;     // Zernike B (n x n) or spline x,y,y2 (3 x n)
    int32_t fit_order, enabled;
    int32_t zap_filter;     // 0 for 'm' and 'r' (raytrace.cpp:747-751)
    int32_t aux2, aux2_n;   // VSHOT points (x, y, x slope, y slope) and their count;
                            // finite element: aux is the flattened node tree (see st_scene.cpp)
    int32_t pad1;
    double roi2;            // VSHOT region of interest, squared radius (m^2)
};

struct DevStage {
    double org[3];
    double R[9];            // global -> stage
    int32_t elem0, nelem;   // slice of the element table
    int32_t virt, multihit, tracethrough;
    int32_t bvh0, nbvh;     // slice of the BVH node table (0 nodes: brute force)
    int32_t always0, nalways;  // with a BVH: slice of the unbounded-element list, tested before the walk
    int32_t generic;        // 1: an element needs the torus or a Newton surface other than flat;
                            // 2: and one is a finite-element surface
};

// Doubles per inverse-CDF segment of the 'd' sunshape (sunshape_icdf).
constexpr int kIcdfStride = 8;

// Sun (coretrace/types.h:195-221) with the sun frame and sun-plane box.
struct DevSun {
    double origin[3];
    double R[9];            // stage -> sun (the sun's RRefToLoc)
    double pos_sun_stage[3];
    double minx, maxx, miny, maxy;
    double sigma, max_angle, max_intensity;
    int32_t shape, point_source;
    int32_t tab, tab_n;     // mirrored table (angles then intensities) in aux
    int32_t tab_ascending;  // 1: angles strictly ascending (binary search is exact)
    int32_t env, env_n;     // upper envelope of intensity/max per polar-angle bin, times 2^32 (aux)
    int32_t icdf, icdf_n;   // inverse-CDF segments of the 'd' polar-angle law in aux (0 segments: none)
    double env_scale;       // bins per mrad
};

// Uniform grid over the sun-plane box listing, per cell, the stage-1
// elements whose sun-plane footprint overlaps it.
struct DevSunGrid {
    double x0, y0, inv_dx, inv_dy;
    int32_t nx, ny;
    int32_t cell0, list0;   // offsets into the cell-start and list buffers
};

// Sun-plane footprint of a stage-1 element's hit box: the convex hull of the
// box projected along the sun direction (a box projects to at most a
// hexagon), as 6 outward half-planes n.p <= off (float, p relative to the
// double centre, padded; unused edges are 0.p <= 1).  A sun ray can only hit
// the element inside one of its footprints.  kind: 0 test the half-planes,
// 1 contains every point, 2 contains none; more: the element's second
// footprint (the next record) is to be tested too.
struct DevFoot {
    double cx, cy;
    float e[6][3];          // nx, ny, off
    int32_t kind, more;
};

// BVH node in depth-first order (an inner node's left child is the next
// node): float box padded outward, the index to continue at when the box is
// missed or the subtree is done, and the leaf item range.
struct DevBvhNode {
    float lo[3];
    int32_t skip;           // next node after this subtree (node count = end)
    float hi[3];
    int32_t leaf;           // 0: inner node; else (first item << 3) | count (1..7)
};

// Beam list of a stage-1 element: the elements a ray leaving it can reach.
// A ray that starts within radius sqrt(r2) of c and whose direction lies
// within the cone (axis a, cos of the half-angle cosa) can only hit items
// of the list; any other ray walks the BVH.  The axis is the mirror
// reflection of the sun direction at the element's vertex; the half-angle
// covers the sunshape, the surface's curvature and its slope and
// specularity errors.
struct DevBeam {
    double c[3];            // stage frame
    float r2;
    float a[3];
    float cosa;
    int32_t first, count;   // slice of the beam item table (count < 0: no list)
    int32_t pad;
};

// A beam list entry: one hit box of element j (stage frame, float, padded
// outward like the BVH's boxes).
struct DevBeamItem {
    float lo[3];
    int32_t j;
    float hi[3];
    int32_t box;            // which of the element's boxes (0 near, 1 far)
};

// Output record in stage frame (coretrace/types.h:242-249 fields).
struct DevRecord {
    double pos[3];
    double cos[3];
    int32_t element;        // 1-based, negative = absorbed, 0 = miss
    uint32_t raynum;
};

// Ray handed from one stage to the next (coretrace/raytrace.cpp:100-111).
struct DevRay {
    double pos[3];
    double cos[3];
    uint32_t raynum;
    uint32_t src;           // 1-based stage-1 element that sent the ray on (0: untagged)
    uint64_t sun_index;
};

struct FluxSpec {
    int32_t active, stage, elem;    // 0-based stage, 1-based element
    int32_t nx, ny, absorbed_only, kind;  // kind 0 flat, 1 cylinder, 2 partial cylinder
    int32_t pad;
    double xmin, ymin, bx, by, radius;
    double org[3];
    double R[9];            // stage -> element
};

}  // namespace st
