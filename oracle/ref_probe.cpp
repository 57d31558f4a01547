// ORACLE — test infrastructure only.  A Python module (built into
// oracle/_ref/ by oracle/Makefile.ref, in a container that holds the
// reference sources; it is git-ignored, and no GPU test, smoke run or bench
// loads it: only the CPU tests of tests/test_oracle_ref.py do)
// that calls single functions of the reference's own Embree-free headers, so
// that tests can compare the CPU restatement in oracle/ with them:
//   intersect.h:71-90      Moller-Trumbore (u, v, t)
//   material.h:257-489     bsdf, sample_bsdf, bsdf_pdf, get_ltc_matrix
//   camera.h:48-320        sample_primary (pinhole and fisheye), unproject,
//                          d_unproject, project_local, d_project_local,
//                          project, d_project, in_screen
//   aabb.cpp:98-311        AABB::intersect(Cone)
//   line_clip.h            clip_line
//   vector.h:690-733       intersect_jacobian, normalize_jacobian
//   distribution.h:7-60    Distribution1D::sample_discrete
//   sample.h:6-21          sample_triangle
// The rest of the reference (scene.h, edge_tree.cpp, pathtrace.h,
// delta_ray.cpp) includes Embree 3, which this image does not have, so it is
// not built.
#include "vector.h"
#include "intersect.h"
#include "material.h"
#include "camera.h"
#include "aabb.h"
#include "line_clip.h"
#include "distribution.h"
#include "sample.h"

#include <pybind11/pybind11.h>
#include <pybind11/numpy.h>
#include <pybind11/stl.h>

#include <vector>

namespace py = pybind11;

namespace {

using FArr = py::array_t<float, py::array::c_style | py::array::forcecast>;

Vector3f v3(const std::vector<float> &a) { return make_vector3(a[0], a[1], a[2]); }
std::vector<float> vec(const Vector3f &a) { return {a[0], a[1], a[2]}; }

Material make_material(FArr diffuse, FArr specular, FArr roughness, FArr uv_d, FArr uv_s, FArr uv_r,
                       bool two_sided) {
    return Material{diffuse, specular, roughness, uv_d, uv_s, uv_r, two_sided};
}

// sp = pos[3], geom_normal[3], frame[9] (rows s, t, n), uv[2]
SurfacePoint make_sp(const std::vector<float> &sp) {
    SurfacePoint p;
    p.position = make_vector3(sp[0], sp[1], sp[2]);
    p.geom_normal = make_vector3(sp[3], sp[4], sp[5]);
    for (int i = 0; i < 3; i++) p.shading_frame[i] = make_vector3(sp[6 + 3 * i], sp[7 + 3 * i], sp[8 + 3 * i]);
    p.uv = make_vector2(sp[15], sp[16]);
    return p;
}

Camera make_camera(FArr c2w, FArr w2c, FArr s2c, FArr c2s, float fov, float aspect, float clip_near,
                   bool fisheye = false) {
    return Camera{c2w, w2c, s2c, c2s, fov, aspect, clip_near, fisheye};
}

}  // namespace

PYBIND11_MODULE(refprobe, m) {
    m.def("mt", [](std::vector<float> v0, std::vector<float> v1, std::vector<float> v2, std::vector<float> org,
                   std::vector<float> dir) {
        Ray r{v3(org), v3(dir)};
        auto uvt = intersect(v3(v0), v3(v1), v3(v2), r);
        return std::vector<float>{std::get<0>(uvt), std::get<1>(uvt), std::get<2>(uvt)};
    });
    m.def("bsdf", [](FArr d, FArr s, FArr r, FArr ud, FArr us, FArr ur, bool two_sided, std::vector<float> sp,
                     std::vector<float> wi, std::vector<float> wo) {
        Material mat = make_material(d, s, r, ud, us, ur, two_sided);
        return vec(bsdf<float>(mat, make_sp(sp), v3(wi), v3(wo)));
    });
    m.def("bsdf_pdf", [](FArr d, FArr s, FArr r, FArr ud, FArr us, FArr ur, bool two_sided, std::vector<float> sp,
                         std::vector<float> wi, std::vector<float> wo) {
        Material mat = make_material(d, s, r, ud, us, ur, two_sided);
        return bsdf_pdf<float>(mat, make_sp(sp), v3(wi), v3(wo));
    });
    m.def("sample_bsdf", [](FArr d, FArr s, FArr r, FArr ud, FArr us, FArr ur, bool two_sided,
                            std::vector<float> sp, std::vector<float> wi, std::vector<float> bs) {
        Material mat = make_material(d, s, r, ud, us, ur, two_sided);
        BSDFSample b{make_vector2(bs[0], bs[1]), bs[2]};
        auto res = sample_bsdf<float>(mat, make_sp(sp), v3(wi), b);
        std::vector<float> out = vec(res.first);
        out.push_back(res.second ? 1.f : 0.f);
        return out;
    });
    m.def("ltc", [](FArr d, FArr s, FArr r, FArr ud, FArr us, FArr ur, bool two_sided, std::vector<float> sp,
                    std::vector<float> wi) {
        Material mat = make_material(d, s, r, ud, us, ur, two_sided);
        Matrix3x3f M = get_ltc_matrix(mat, make_sp(sp), v3(wi));
        std::vector<float> out;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) out.push_back(M(i, j));
        return out;
    });
    m.def("sample_primary", [](FArr c2w, FArr w2c, FArr s2c, FArr c2s, float fov, float aspect, float clip_near,
                               std::vector<float> screen) {
        Camera cam = make_camera(c2w, w2c, s2c, c2s, fov, aspect, clip_near);
        TCamera<float> tc(cam);
        auto ray = sample_primary(tc, make_vector2(screen[0], screen[1]));
        std::vector<float> out = vec(ray.org);
        auto d = vec(ray.dir);
        out.insert(out.end(), d.begin(), d.end());
        return out;
    });
    m.def("sample_primary_fisheye", [](FArr c2w, FArr w2c, FArr s2c, FArr c2s, float fov, float aspect,
                                       float clip_near, std::vector<float> screen) {
        Camera cam = make_camera(c2w, w2c, s2c, c2s, fov, aspect, clip_near, true);
        TCamera<float> tc(cam);
        auto ray = sample_primary(tc, make_vector2(screen[0], screen[1]));
        std::vector<float> out = vec(ray.org);
        auto d = vec(ray.dir);
        out.insert(out.end(), d.begin(), d.end());
        return out;
    });
    // unproject, d_unproject (x, y), project_local, d_project_local, in_screen
    m.def("camera_functions", [](FArr c2w, FArr w2c, FArr s2c, FArr c2s, float fov, float aspect, float clip_near,
                                 bool fisheye, std::vector<float> screen, std::vector<float> local,
                                 std::vector<float> dxy) {
        Camera cam = make_camera(c2w, w2c, s2c, c2s, fov, aspect, clip_near, fisheye);
        auto sp = make_vector2(screen[0], screen[1]);
        auto u = unproject(cam, sp);
        auto du = d_unproject(cam, sp);
        auto p = project_local(cam, v3(local));
        auto dl = d_project_local(cam, v3(local), dxy[0], dxy[1]);
        std::vector<float> out = vec(u);
        for (auto &x : vec(du.first)) out.push_back(x);
        for (auto &x : vec(du.second)) out.push_back(x);
        out.push_back(p[0]);
        out.push_back(p[1]);
        for (auto &x : vec(dl)) out.push_back(x);
        out.push_back(in_screen(cam, sp) ? 1.f : 0.f);
        return out;
    });
    m.def("project", [](FArr c2w, FArr w2c, FArr s2c, FArr c2s, float fov, float aspect, float clip_near,
                        std::vector<float> p0, std::vector<float> p1) {
        Camera cam = make_camera(c2w, w2c, s2c, c2s, fov, aspect, clip_near);
        auto r = project(cam, v3(p0), v3(p1));
        if (!r) return std::vector<float>{};
        return std::vector<float>{r->first[0], r->first[1], r->second[0], r->second[1]};
    });
    m.def("d_project", [](FArr c2w, FArr w2c, FArr s2c, FArr c2s, float fov, float aspect, float clip_near,
                          std::vector<float> p0, std::vector<float> p1, std::vector<float> dxy) {
        Camera cam = make_camera(c2w, w2c, s2c, c2s, fov, aspect, clip_near);
        Matrix4x4f d_c2s, d_w2c;
        d_project(cam, v3(p0), v3(p1), dxy[0], dxy[1], dxy[2], dxy[3], d_c2s, d_w2c);
        std::vector<float> out;
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) out.push_back(d_w2c(i, j));
        return out;
    });
    m.def("cone_box", [](std::vector<float> lo, std::vector<float> hi, std::vector<float> org, std::vector<float> dir,
                         float cos_angle) {
        AABB box(v3(lo), v3(hi));
        Cone cone{Ray{v3(org), v3(dir)}, cos_angle};
        return box.intersect(cone);
    });
    m.def("clip_line", [](std::vector<float> a, std::vector<float> b) {
        auto r = clip_line(make_vector2(a[0], a[1]), make_vector2(b[0], b[1]));
        if (!r) return std::vector<float>{};
        return std::vector<float>{r->first[0], r->first[1], r->second[0], r->second[1]};
    });
    m.def("jacobians", [](std::vector<float> x, std::vector<float> dir, std::vector<float> p, std::vector<float> n) {
        Matrix3x3f nj = normalize_jacobian(v3(x));
        Matrix3x3f ij = intersect_jacobian(v3(dir), v3(p), v3(n));
        std::vector<float> out;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) out.push_back(nj(i, j));
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) out.push_back(ij(i, j));
        return out;
    });
    m.def("sample_discrete", [](std::vector<float> f, float u) {
        Distribution1D d(f.data(), (int)f.size());
        float pmf = 0.f;
        int i = d.sample_discrete(u, &pmf);
        return std::make_pair(i, pmf);
    });
    m.def("sample_triangle", [](std::vector<float> v0, std::vector<float> v1, std::vector<float> v2,
                                std::vector<float> u) {
        SurfacePoint sp = sample_triangle(v3(v0), v3(v1), v3(v2), make_vector2(u[0], u[1]));
        std::vector<float> out = vec(sp.position);
        auto n = vec(sp.geom_normal);
        out.insert(out.end(), n.begin(), n.end());
        return out;
    });
}
