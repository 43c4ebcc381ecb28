// Host scene model and scene compiler of the HIP backend.
//
// The model mirrors what the stapi.h setters describe (coretrace/types.h:
// 85-343); compile() does the work of the reference's InitGeometries,
// TranslateSurfaceParams and AperturePlane (coretrace/input.cpp:153-366,
// apertureplane.cpp:57-491) and of SunToPrimaryStage
// (suntoprimarystage.cpp:59-293), and lays the result out as the flat
// device tables of st_types.h plus the cull structures (sun-plane grid,
// per-stage BVH).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "st_fetree.h"
#include "st_types.h"

namespace st {

struct HFace {
    char dist = 'g';
    int optnum = 1, apgr = 0, order = 0;
    double refr[4] = {0, 0, 0, 0};
    double ab12[4] = {0, 0, 0, 0};
    double reflectivity = 0, transmissivity = 0, slope = 0, spec = 0;
    bool use_refl = false, use_trans = false;
    std::vector<double> refl_ang, refl_val, trans_ang, trans_val;
};

struct HOptic {
    std::string name;
    HFace front, back;
};

struct HElement {
    bool enabled = true;
    double origin[3] = {0, 0, 0}, aim[3] = {0, 0, 0}, zrot = 0;
    char shape = ' ';
    double p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    char surf = ' ';
    int stype = 0;                       // reference surface type code
    double kappa = 0, alpha[5] = {0, 0, 0, 0, 0};
    double vcx = 0, vcy = 0, ann_r = 0, cross_r = 0, cone_half = 0, curv_rev = 0;
    int fit_order = 0;
    std::vector<double> zern;            // (order+1)^2 row-major B(i,j)
    std::vector<double> poly;            // .ply coefficients
    std::vector<double> sx, sy, sy2;     // cubic spline
    std::vector<double> vshot;           // VSHOT points: x, y, x slope, y slope
    std::vector<double> fe_nodes;        // finite-element nodes: x, y, z
    st_fe::Tree fe_tree;                 // their spatial hash tree
    int interaction = 0;
    std::string optic_name;
    // filled by compile()
    double euler[3] = {0, 0, 0};
    double R[9] = {0};
    double zap = 0;
};

struct HStage {
    bool multihit = true, virt = false, tracethrough = false;
    double origin[3] = {0, 0, 0}, aim[3] = {0, 0, 0}, zrot = 0;
    std::vector<HElement> elements;
    double euler[3] = {0, 0, 0};
    double R[9] = {0};
};

struct HSun {
    char shape = ' ';
    double sigma = 0;
    bool point_source = false;
    std::vector<double> angle, intensity;  // mirrored table
    double max_angle = 0, max_intensity = 0;
    double origin[3] = {0, 0, 0};
    // filled by compile()
    double minx = 0, maxx = 0, miny = 0, maxy = 0;
};

// Device-table image of a compiled scene, still on the host.
struct CompiledScene {
    std::vector<DevElem> elems;
    std::vector<DevFace> faces;
    std::vector<double> aux;
    std::vector<DevStage> stages;
    std::vector<DevBvhNode> bvh;
    std::vector<int32_t> bvh_items;
    std::vector<int32_t> always_items;  // per stage with a BVH: elements no box bounds
    std::vector<int32_t> grid_cells, grid_items;
    // stage-1 elements no footprint bounds (torus, finite element): tested
    // for every sun ray, kept out of the grid's cell lists
    std::vector<int32_t> grid_always;
    std::vector<DevFoot> foot;  // 2 per stage-1 element, one per hit box
    std::vector<double> ebox;   // per element: 2 element-frame hit boxes lo[3], hi[3]
    std::vector<int8_t> enbox;  // per element: boxes that bound its hits (0: none)
    // stage-1 hit boxes in stage frame, as the BVH holds them (j = element)
    std::vector<DevBeamItem> s0_boxes;
    std::vector<DevBeam> beam;            // per stage-1 element (empty: no lists)
    std::vector<DevBeamItem> beam_items;
    DevSunGrid grid{};
    bool use_grid = false;
    DevSun sun{};
};

struct SceneModel {
    HSun sun;
    std::vector<HOptic> optics;
    std::vector<HStage> stages;
};

// Parameter translation at st_element_surface_params time
// (coretrace/input.cpp:288-366); message on failure.
bool translate_surface_params(HElement &e, const double params[8], std::string &msg);
// Surface file read at st_element_surface_file time (.mon, .csi, .ply;
// coretrace/input.cpp:368-509).
bool read_surface_file(HElement &e, const char *file, std::string &msg);

// InitGeometries + SunToPrimaryStage + device-table layout.  Returns false
// with the reference's message on invalid input.
bool compile_scene(SceneModel &m, bool use_cull, CompiledScene &out, std::string &msg);
// Sun frame, sun-plane box and sun-plane grid; after compile_scene.
bool compile_sun(SceneModel &m, CompiledScene &out, std::string &msg);

// Shared math (same formulas as coretrace/mathproc.cpp).
void euler_matrix(const double e[3], double R[9]);
bool euler_from_aim(const double org[3], const double aim[3], double zrot, double e[3]);

}  // namespace st
