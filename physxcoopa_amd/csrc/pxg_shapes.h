// Device helpers over the packed shape table shared by the step kernels and the cloth and hair
// kernels: a shape's decoded geometry, the hull and mesh tables, world bounds and CCD extent of
// any shape type, and liveness.
#pragma once
#include "pxg_state.h"
#include "pxg_mesh.h"

namespace pxg {

using namespace pxd;

__device__ __forceinline__ ShapeGeom load_shape(const DevState& d, uint32_t i, uint4 info) {
    ShapeGeom s;
    s.type = int(info.y & 15u);
    s.caxis = int((info.y >> 8) & 3u);
    float4 lc = d.s_lc[i];
    s.lc = mk(lc.x, lc.y, lc.z);
    s.radius = lc.w;
    s.lr = ldq(d.s_lr[i]);
    float4 dim = d.s_dim[i];
    s.he = mk(dim.x, dim.y, dim.z);
    s.cr = dim.w;
    s.chh = d.s_mat[i].x;
    s.hull = __float_as_uint(dim.x);
    return s;
}
__device__ __forceinline__ HullTab hull_tab(const DevState& d) {
    HullTab t;
    t.meta = d.h_meta; t.meta2 = d.h_meta2; t.lb = d.h_lb; t.cen = d.h_cen; t.v = d.h_v; t.pl = d.h_pl;
    t.face = d.h_face; t.fv = d.h_fv; t.e = d.h_e;
    return t;
}
__device__ __forceinline__ MeshTab mesh_tab(const DevState& d) {
    MeshTab t;
    t.meta = d.me_meta; t.lb = d.me_lb; t.v = d.me_v; t.tri = d.me_tri; t.n = d.me_n; t.adj = d.me_adj;
    t.node_lo = d.me_node_lo; t.node_hi = d.me_node_hi; t.node_tri = d.me_node_tri; t.order = d.me_order;
    return t;
}
// tight world AABB and CCD extent of any shape type
__device__ __forceinline__ void shape_bounds(const DevState& d, const ShapeGeom& s, v3 pos, quat rot, v3& mn, v3& mx) {
    if (s.type == SH_HULL) hull_bounds(hull_tab(d), s, pos, rot, mn, mx);
    else if (s.type == SH_MESH) mesh_bounds(mesh_tab(d), s, pos, rot, mn, mx);
    else world_bounds(s, pos, rot, mn, mx);
}
__device__ __forceinline__ float shape_min_extent(const DevState& d, const ShapeGeom& s) {
    if (s.type == SH_MESH) return 3.402823466e+38f;  // world.h:1416: a mesh never gates speculation
    return s.type == SH_HULL ? hull_min_extent(hull_tab(d), s) : min_extent(s);
}
__device__ __forceinline__ bool shape_live(const DevState& d, uint4 info) {
    return (info.y & SF_ENABLED) && (d.b_flags[info.x] & BF_ALIVE);
}


}  // namespace pxg
