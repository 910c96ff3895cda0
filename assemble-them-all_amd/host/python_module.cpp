// pybind11 module atac._redmax: the C++ contact classes (Body, BodySDFObj,
// ForceGeneralSDFContact, BodyBVHObj, ForceGeneralBVHContact, BatchedBDF1) for
// Python, as the reference's redmax_py exposes its
// simulator (simulation/redmax/python_interface.cpp:34-303). Arrays are numpy FP64.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "Body/Body.h"
#include "Body/BodyBVHObj.h"
#include "Body/BodySDFObj.h"
#include "Body/ContactPoints.h"
#include "Simulation/BatchedBDF1.h"
#include "Force/ForceGeneralBVHContact.h"
#include "Force/ForceGeneralSDFContact.h"

namespace py = pybind11;
using namespace redmax;

namespace {

using arr = py::array_t<double, py::array::c_style | py::array::forcecast>;

std::vector<Vector3> points_of(const arr& a) {
    if (a.ndim() != 2 || a.shape(1) != 3) throw std::invalid_argument("contact points must have shape (n, 3)");
    std::vector<Vector3> v((size_t)a.shape(0));
    auto r = a.unchecked<2>();
    for (py::ssize_t i = 0; i < a.shape(0); ++i) v[(size_t)i] = make_vector3(r(i, 0), r(i, 1), r(i, 2));
    return v;
}

Vector3 vec3(const arr& a) {
    if (a.size() != 3) throw std::invalid_argument("expected 3 values");
    const double* p = a.data();
    return make_vector3(p[0], p[1], p[2]);
}

py::array_t<double> mat_out(const MatrixX& m) {
    py::array_t<double> out({m.rows(), m.cols()});
    auto w = out.mutable_unchecked<2>();
    for (int r = 0; r < m.rows(); ++r)
        for (int c = 0; c < m.cols(); ++c) w(r, c) = m(r, c);
    return out;
}

py::array_t<double> vec_out(const VectorX& v) {
    py::array_t<double> out(v.size());
    auto w = out.mutable_unchecked<1>();
    for (int i = 0; i < v.size(); ++i) w(i) = v(i);
    return out;
}

template <int R, int C>
py::array_t<double> fixed_out(const FixedMatrix<R, C>& m) {
    py::array_t<double> out({R, C});
    auto w = out.mutable_unchecked<2>();
    for (int r = 0; r < R; ++r)
        for (int c = 0; c < C; ++c) w(r, c) = m(r, c);
    return out;
}

Matrix4 m4(const arr& a) {
    if (a.size() != 16) throw std::invalid_argument("expected a 4x4 matrix");
    Matrix4 m;
    const double* p = a.data();
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) m(r, c) = p[4 * r + c];
    return m;
}

Vector6 v6(const arr& a) {
    if (a.size() != 6) throw std::invalid_argument("expected 6 values");
    Vector6 v;
    for (int k = 0; k < 6; ++k) v(k) = a.data()[k];
    return v;
}

sdfgen::Array3<dtype> grid_of(const arr& values) {
    if (values.ndim() != 3) throw std::invalid_argument("grid values must have shape (nk, nj, ni)");
    sdfgen::Array3<dtype> g;
    g.resize((int)values.shape(2), (int)values.shape(1), (int)values.shape(0));
    std::copy(values.data(), values.data() + values.size(), g.a.begin());
    return g;
}

}  // namespace

PYBIND11_MODULE(_redmax, m) {
    m.doc() = "GPU-backed SDF and BVH contact forces with the reference simulator's C++ interface";

    m.def(
        "precompute_contact_points",
        [](const arr& V, py::array_t<int64_t, py::array::c_style | py::array::forcecast> F, bool adaptive_sample) {
            const std::vector<Vector3> v = points_of(V);
            if (F.ndim() != 2 || F.shape(1) != 3) throw std::invalid_argument("faces must have shape (m, 3)");
            std::vector<std::array<int, 3>> f((size_t)F.shape(0));
            auto r = F.unchecked<2>();
            for (py::ssize_t t = 0; t < F.shape(0); ++t)
                for (int k = 0; k < 3; ++k) {
                    if (r(t, k) < 0 || r(t, k) >= (int64_t)v.size()) throw std::invalid_argument("face index out of range");
                    f[(size_t)t][(size_t)k] = (int)r(t, k);
                }
            int rate = 1;
            const std::vector<Vector3> p = precompute_contact_points(v, f, adaptive_sample, &rate);
            py::array_t<double> out({(py::ssize_t)p.size(), (py::ssize_t)3});
            auto w = out.mutable_unchecked<2>();
            for (size_t i = 0; i < p.size(); ++i)
                for (int k = 0; k < 3; ++k) w((py::ssize_t)i, k) = p[i](k);
            return py::make_tuple(out, rate);
        },
        py::arg("V"), py::arg("F"), py::arg("adaptive_sample") = false,
        "BodyMeshObj::precompute_contact_points (BodyMeshObj.cpp:184-238): (points (n, 3), sample_rate).");

    py::class_<Body>(m, "Body")
        .def(py::init([](std::string name, const arr& pts, int index0) {
                 return new Body(nullptr, name, points_of(pts), index0);
             }),
             py::arg("name"), py::arg("contact_points"), py::arg("index0") = 0)
        .def_readwrite("name", &Body::_name)
        .def_readwrite("contact_scale", &Body::_contact_scale)
        .def_readwrite("p6_column", &Body::_p6_column)
        .def_readonly("index", &Body::_index)
        .def_property("E_0i", [](const Body& b) { return fixed_out(b._E_0i); },
                      [](Body& b, const arr& a) { b._E_0i = m4(a); })
        .def_property("phi", [](const Body& b) { return fixed_out(b._phi); },
                      [](Body& b, const arr& a) { b._phi = v6(a); })
        .def("get_contact_points", [](const Body& b) {
            py::array_t<double> out({(py::ssize_t)b._contact_points.size(), (py::ssize_t)3});
            auto w = out.mutable_unchecked<2>();
            for (size_t i = 0; i < b._contact_points.size(); ++i)
                for (int k = 0; k < 3; ++k) w(i, k) = b._contact_points[i](k);
            return out;
        })
        .def("get_contact_bodies", &Body::get_contact_bodies)
        .def("clear_contact_bodies", &Body::clear_contact_bodies);

    py::class_<BodySDFObj, Body>(m, "BodySDFObj")
        .def(py::init([](std::string name, const arr& values, const arr& min_box, const arr& max_box, const arr& dxyz,
                         double col_th, const arr& pts, int index0) {
                 const double* d = dxyz.data();
                 return new BodySDFObj(nullptr, name, grid_of(values), vec3(min_box), vec3(max_box), d[0], d[1], d[2],
                                       col_th, points_of(pts), index0);
             }),
             py::arg("name"), py::arg("values"), py::arg("min_box"), py::arg("max_box"), py::arg("dxyz"),
             py::arg("col_th"), py::arg("contact_points"), py::arg("index0") = 0)
        .def_static("from_sdf_file", [](std::string name, std::string path, double col_th, const arr& pts, int index0) {
                        sdfgen::Array3<dtype> g;
                        Vector3 mn, mx;
                        dtype dx, dy, dz;
                        if (!BodySDFObj::load_SDF_file(path, g, mn, mx, dx, dy, dz))
                            throw std::runtime_error("cannot read " + path);
                        return new BodySDFObj(nullptr, name, g, mn, mx, dx, dy, dz, col_th, points_of(pts), index0);
                    },
                    py::arg("name"), py::arg("path"), py::arg("col_th"), py::arg("contact_points"), py::arg("index0") = 0)
        .def_static("from_mesh", [](std::string name, const arr& V, const py::array_t<int64_t, py::array::c_style | py::array::forcecast>& F,
                                    double dx, int res, double col_th, const arr& pts, int index0) {
                        if (V.ndim() != 2 || V.shape(1) != 3 || F.ndim() != 2 || F.shape(1) != 3)
                            throw std::runtime_error("from_mesh: V is n x 3, F is m x 3");
                        const std::vector<Vector3> vs = points_of(V);
                        std::vector<std::array<int, 3>> fs((size_t)F.shape(0));
                        const int64_t* fp = F.data();
                        for (size_t t = 0; t < fs.size(); ++t)
                            for (int a = 0; a < 3; ++a) fs[t][a] = (int)fp[3 * t + a];
                        sdfgen::Array3<dtype> g;
                        Vector3 mn, mx;
                        dtype dx_, dy_, dz_;
                        BodySDFObj::precompute_SDF(vs, fs, dx, res, g, mn, mx, dx_, dy_, dz_);
                        return new BodySDFObj(nullptr, name, g, mn, mx, dx_, dy_, dz_, col_th, points_of(pts), index0);
                    },
                    py::arg("name"), py::arg("V"), py::arg("F"), py::arg("dx"), py::arg("res"), py::arg("col_th"),
                    py::arg("contact_points"), py::arg("index0") = 0)
        .def("save_sdf", &BodySDFObj::save_SDF_file)
        .def_readwrite("col_th", &BodySDFObj::_col_th)
        .def_property_readonly("min_box", [](const BodySDFObj& b) { return fixed_out(b._min_box); })
        .def_property_readonly("max_box", [](const BodySDFObj& b) { return fixed_out(b._max_box); })
        .def_property_readonly("dxyz", [](const BodySDFObj& b) { return std::vector<double>{b._dx, b._dy, b._dz}; })
        .def_property_readonly("shape", [](const BodySDFObj& b) { return std::vector<int>{b._SDF.ni, b._SDF.nj, b._SDF.nk}; })
        .def("values", [](const BodySDFObj& b) {
            py::array_t<double> out({(py::ssize_t)b._SDF.nk, (py::ssize_t)b._SDF.nj, (py::ssize_t)b._SDF.ni});
            std::copy(b._SDF.a.begin(), b._SDF.a.end(), out.mutable_data());
            return out;
        })
        .def("distance", [](BodySDFObj& b, const arr& xw) { return b.distance(vec3(xw)); })
        .def("collision", [](BodySDFObj& b, const arr& xw, const arr& xw_dot, bool derivs) {
            dtype d, ddot;
            Vector3 n, tdot, xi2;
            py::dict out;
            if (!derivs) {
                b.collision(vec3(xw), vec3(xw_dot), d, n, ddot, tdot, xi2);
            } else {
                RowVector3 dd_dxw, dddot_dxw, dddot_dxwdot;
                RowVector6 dd_dq2, dddot_dq2, dddot_dphi2;
                Matrix3 dn_dxw, dtdot_dxw, dtdot_dxwdot, dxi2_dxw;
                Matrix36 dn_dq2, dtdot_dq2, dtdot_dphi2, dxi2_dq2;
                b.collision(vec3(xw), vec3(xw_dot), d, n, ddot, tdot, xi2, dd_dxw, dd_dq2, dn_dxw, dn_dq2, dddot_dxw,
                            dddot_dxwdot, dddot_dq2, dddot_dphi2, dtdot_dxw, dtdot_dxwdot, dtdot_dq2, dtdot_dphi2,
                            dxi2_dxw, dxi2_dq2);
                out["dd_dxw"] = fixed_out(dd_dxw); out["dd_dq2"] = fixed_out(dd_dq2);
                out["dn_dxw"] = fixed_out(dn_dxw); out["dn_dq2"] = fixed_out(dn_dq2);
                out["dddot_dxw"] = fixed_out(dddot_dxw); out["dddot_dxwdot"] = fixed_out(dddot_dxwdot);
                out["dddot_dq2"] = fixed_out(dddot_dq2); out["dddot_dphi2"] = fixed_out(dddot_dphi2);
                out["dtdot_dxw"] = fixed_out(dtdot_dxw); out["dtdot_dxwdot"] = fixed_out(dtdot_dxwdot);
                out["dtdot_dq2"] = fixed_out(dtdot_dq2); out["dtdot_dphi2"] = fixed_out(dtdot_dphi2);
                out["dxi2_dxw"] = fixed_out(dxi2_dxw); out["dxi2_dq2"] = fixed_out(dxi2_dq2);
            }
            out["d"] = d; out["ddot"] = ddot;
            out["n"] = fixed_out(n); out["tdot"] = fixed_out(tdot); out["xi2"] = fixed_out(xi2);
            return out;
        }, py::arg("xw"), py::arg("xw_dot"), py::arg("derivs") = false);

    py::class_<ForceGeneralSDFContact>(m, "ForceGeneralSDFContact")
        .def(py::init([](Body* a, Body* b, double kn, double kt, double mu, double damping) {
                 return new ForceGeneralSDFContact(nullptr, a, b, kn, kt, mu, damping);
             }),
             py::arg("contact_body"), py::arg("SDF_body"), py::arg("kn") = 1.0, py::arg("kt") = 0.0,
             py::arg("mu") = 0.0, py::arg("damping") = 0.0, py::keep_alive<1, 2>(), py::keep_alive<1, 3>())
        .def("set_stiffness", &ForceGeneralSDFContact::set_stiffness)
        .def("set_friction", &ForceGeneralSDFContact::set_friction)
        .def("set_damping", &ForceGeneralSDFContact::set_damping)
        .def("set_scale", &ForceGeneralSDFContact::set_scale)
        .def_readonly("kn", &ForceGeneralSDFContact::_kn)
        .def_readonly("kt", &ForceGeneralSDFContact::_kt)
        .def_readonly("mu", &ForceGeneralSDFContact::_mu)
        .def_readonly("damping", &ForceGeneralSDFContact::_damping)
        .def_readonly("scale", &ForceGeneralSDFContact::_scale)
        .def("computeForce", [](ForceGeneralSDFContact& f, int ndof_m) {
            VectorX fm = VectorX::Zero(ndof_m), fr;
            f.computeForce(fm, fr);
            return vec_out(fm);
        }, py::arg("ndof_m"))
        .def("computeForceWithDerivative", [](ForceGeneralSDFContact& f, int ndof_m) {
            VectorX fm = VectorX::Zero(ndof_m), fr;
            MatrixX Km = MatrixX::Zero(ndof_m, ndof_m), Dm = MatrixX::Zero(ndof_m, ndof_m), Kr, Dr;
            f.computeForceWithDerivative(fm, fr, Km, Dm, Kr, Dr);
            return py::make_tuple(vec_out(fm), mat_out(Km), mat_out(Dm));
        }, py::arg("ndof_m"))
        .def("computeForceWithDesignDerivative", [](ForceGeneralSDFContact& f, int ndof_m, int ndof_p) {
            VectorX fm = VectorX::Zero(ndof_m), fr;
            MatrixX Km = MatrixX::Zero(ndof_m, ndof_m), Dm = MatrixX::Zero(ndof_m, ndof_m), Kr, Dr;
            MatrixX dfm_dp = MatrixX::Zero(ndof_m, ndof_p), dfr_dp;
            f.computeForceWithDerivative(fm, fr, Km, Dm, Kr, Dr, dfm_dp, dfr_dp);
            return py::make_tuple(vec_out(fm), mat_out(Km), mat_out(Dm), mat_out(dfm_dp));
        }, py::arg("ndof_m"), py::arg("ndof_p"))
        .def("test_derivatives", &ForceGeneralSDFContact::test_derivatives)
        .def("test_derivatives_runtime", &ForceGeneralSDFContact::test_derivatives_runtime_error)
        .def("test_design_derivatives_runtime", &ForceGeneralSDFContact::test_design_derivatives_runtime)
        .def("last_active_count", &ForceGeneralSDFContact::last_active_count)
        .def("time_report_us", [](ForceGeneralSDFContact& f) {
            return py::make_tuple(f._time.get("collision detection"), f._time.get("compute force"));
        });

    py::class_<BodyBVHObj, Body>(m, "BodyBVHObj")
        .def(py::init([](std::string name, const arr& V, const py::array_t<int64_t, py::array::c_style | py::array::forcecast>& F,
                         double col_th, const arr& pts, int index0) {
                 if (F.ndim() != 2 || F.shape(1) != 3) throw std::invalid_argument("triangles must have shape (m, 3)");
                 std::vector<std::array<int, 3>> fs((size_t)F.shape(0));
                 auto r = F.unchecked<2>();
                 for (py::ssize_t i = 0; i < F.shape(0); ++i) fs[(size_t)i] = {(int)r(i, 0), (int)r(i, 1), (int)r(i, 2)};
                 return new BodyBVHObj(nullptr, name, points_of(V), fs, col_th, points_of(pts), index0);
             }),
             py::arg("name"), py::arg("V"), py::arg("F"), py::arg("col_th"), py::arg("contact_points"),
             py::arg("index0") = 0)
        .def_readwrite("col_th", &BodyBVHObj::_col_th)
        .def("get_AABB", [](const BodyBVHObj& b) {
            auto bb = b.get_AABB();
            return py::make_tuple(fixed_out(bb.first), fixed_out(bb.second));
        })
        .def("distance", [](BodyBVHObj& b, const arr& xw) { dtype d; b.distance(vec3(xw), d); return d; })
        .def("distance_normal", [](BodyBVHObj& b, const arr& xw) {
            dtype d;
            Vector3 g;
            b.distance_normal(vec3(xw), d, g);
            return py::make_tuple(d, fixed_out(g));
        })
        .def("distance_normal_parallel", [](BodyBVHObj& b, const arr& xw) {
            std::vector<dtype> d;
            std::vector<Vector3> g;
            b.distance_normal_parallel(points_of(xw), d, g);
            py::array_t<double> gd({(py::ssize_t)g.size(), (py::ssize_t)3});
            auto w = gd.mutable_unchecked<2>();
            for (size_t i = 0; i < g.size(); ++i)
                for (int k = 0; k < 3; ++k) w(i, k) = g[i](k);
            return py::make_tuple(py::array_t<double>(d.size(), d.data()), gd);
        });

    py::class_<ForceGeneralBVHContact>(m, "ForceGeneralBVHContact")
        .def(py::init([](Body* a, Body* b, double kn, double kt, double mu, double damping) {
                 return new ForceGeneralBVHContact(nullptr, a, b, kn, kt, mu, damping);
             }),
             py::arg("contact_body"), py::arg("BVH_body"), py::arg("kn") = 1.0, py::arg("kt") = 0.0,
             py::arg("mu") = 0.0, py::arg("damping") = 0.0, py::keep_alive<1, 2>(), py::keep_alive<1, 3>())
        .def("set_stiffness", &ForceGeneralBVHContact::set_stiffness)
        .def("set_friction", &ForceGeneralBVHContact::set_friction)
        .def("set_damping", &ForceGeneralBVHContact::set_damping)
        .def("set_scale", &ForceGeneralBVHContact::set_scale)
        .def("computeForce", [](ForceGeneralBVHContact& f, int ndof_m) {
            VectorX fm = VectorX::Zero(ndof_m), fr;
            f.computeForce(fm, fr);
            return vec_out(fm);
        }, py::arg("ndof_m"))
        .def("computeForceWithDerivative", [](ForceGeneralBVHContact& f, int ndof_m) {
            VectorX fm = VectorX::Zero(ndof_m), fr;
            MatrixX Km = MatrixX::Zero(ndof_m, ndof_m), Dm = MatrixX::Zero(ndof_m, ndof_m), Kr, Dr;
            f.computeForceWithDerivative(fm, fr, Km, Dm, Kr, Dr);
            return py::make_tuple(vec_out(fm), mat_out(Km), mat_out(Dm));
        }, py::arg("ndof_m"))
        .def("computeForceWithDesignDerivative", [](ForceGeneralBVHContact& f, int ndof_m, int ndof_p) {
            VectorX fm = VectorX::Zero(ndof_m), fr;
            MatrixX Km = MatrixX::Zero(ndof_m, ndof_m), Dm = MatrixX::Zero(ndof_m, ndof_m), Kr, Dr;
            MatrixX dfm_dp = MatrixX::Zero(ndof_m, ndof_p), dfr_dp;
            f.computeForceWithDerivative(fm, fr, Km, Dm, Kr, Dr, dfm_dp, dfr_dp);
            return py::make_tuple(vec_out(fm), mat_out(Km), mat_out(Dm), mat_out(dfm_dp));
        }, py::arg("ndof_m"), py::arg("ndof_p"))
        .def("last_active_count", &ForceGeneralBVHContact::last_active_count);

    py::class_<BatchedBDF1>(m, "BatchedBDF1",
                            "BDF1 steps of parts on translational joints against fixed parts, many candidates at once")
        .def(py::init([](py::list parts, std::vector<std::pair<int, int>> contacts, double kn, double kt, double mu,
                         double damping, double col_th, double h, std::vector<double> gravity, bool precise) {
                 if (gravity.size() != 3) throw std::invalid_argument("gravity has 3 components");
                 std::vector<BDF1Part> ps;
                 for (auto item : parts) {
                     auto t = item.cast<py::tuple>();
                     if (t.size() != 3) throw std::invalid_argument("a part is (points, BodySDFObj or None, E0)");
                     BDF1Part p;
                     const arr pts = t[0].cast<arr>();
                     if (pts.size() > 0) p.points = points_of(pts);
                     if (!t[1].is_none()) p.grid = t[1].cast<BodySDFObj*>()->device_grid();
                     const arr E0 = t[2].cast<arr>();
                     if (E0.size() != 16) throw std::invalid_argument("E0 must be 4 x 4");
                     std::copy(E0.data(), E0.data() + 16, p.E0.begin());
                     ps.push_back(std::move(p));
                 }
                 BDF1ContactParams cp;
                 cp.kn = kn; cp.kt = kt; cp.mu = mu; cp.damping = damping; cp.col_th = col_th;
                 return new BatchedBDF1(std::move(ps), std::move(contacts), cp, h,
                                        make_vector3(gravity[0], gravity[1], gravity[2]), precise);
             }),
             py::arg("parts"), py::arg("contacts"), py::arg("kn") = 1e6, py::arg("kt") = 1e3, py::arg("mu") = 0.0,
             py::arg("damping") = 0.0, py::arg("col_th") = 0.01, py::arg("h") = 1e-3,
             py::arg("gravity") = std::vector<double>{0.0, 0.0, 0.0}, py::arg("precise") = false,
             py::keep_alive<1, 2>())
        .def("add_candidate", [](BatchedBDF1& s, int move, double mass, const arr& q, const arr& qdot, const arr& u) {
                 return s.add_candidate(move, mass, vec3(q), vec3(qdot), vec3(u));
             }, py::arg("move"), py::arg("mass"), py::arg("q"), py::arg("qdot"), py::arg("u"))
        .def("set_state", [](BatchedBDF1& s, int c, const arr& q, const arr& qdot) { s.set_state(c, vec3(q), vec3(qdot)); })
        .def("set_force", [](BatchedBDF1& s, int c, const arr& u) { s.set_force(c, vec3(u)); })
        .def("step", &BatchedBDF1::step, py::arg("n") = 1, py::call_guard<py::gil_scoped_release>())
        .def("q", [](const BatchedBDF1& s, int c) { return fixed_out(s.q(c)); })
        .def("qdot", [](const BatchedBDF1& s, int c) { return fixed_out(s.qdot(c)); })
        .def("converged", &BatchedBDF1::converged)
        .def("newton_iterations", &BatchedBDF1::newton_iterations)
        .def_property_readonly("n_candidates", &BatchedBDF1::n_candidates)
        .def_property_readonly("n_evaluations", &BatchedBDF1::n_evaluations);
}
