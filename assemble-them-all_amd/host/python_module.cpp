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
#include "Body/MeshObj.h"
#include "Joint/Joint.h"
#include "Simulation/BatchedBDF1.h"
#include "Simulation/Simulation.h"
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

VectorX vec_in(const arr& a) {
    VectorX v((int)a.size());
    for (py::ssize_t i = 0; i < a.size(); ++i) v((int)i) = a.data()[i];
    return v;
}

py::array_t<double> points_out(const std::vector<Vector3>& V, bool rows_are_points) {
    const py::ssize_t n = (py::ssize_t)V.size();
    py::array_t<double> out(rows_are_points ? std::vector<py::ssize_t>{n, 3} : std::vector<py::ssize_t>{3, n});
    auto w = out.mutable_unchecked<2>();
    for (py::ssize_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            if (rows_are_points) w(i, a) = V[(size_t)i](a);
            else w(a, i) = V[(size_t)i](a);
        }
    return out;
}

py::array_t<int64_t> faces_out(const std::vector<std::array<int, 3>>& F) {
    py::array_t<int64_t> out(std::vector<py::ssize_t>{3, (py::ssize_t)F.size()});
    auto w = out.mutable_unchecked<2>();
    for (size_t t = 0; t < F.size(); ++t)
        for (int a = 0; a < 3; ++a) w(a, (py::ssize_t)t) = F[t][a];
    return out;
}

py::array_t<double> row_major(const double* p, std::vector<py::ssize_t> shape) {
    py::array_t<double> out(shape);
    std::copy(p, p + out.size(), out.mutable_data());
    return out;
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
    // throw_error throws the message as a std::string: raise it as RuntimeError(message)
    py::register_exception_translator([](std::exception_ptr p) {
        try {
            if (p) std::rethrow_exception(p);
        } catch (const std::string& msg) {
            PyErr_SetString(PyExc_RuntimeError, msg.c_str());
        }
    });

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
        .def(py::init([](std::string filename, const arr& R, const arr& p, double dx, int res, double col_th,
                         std::string transform_type, double density, const arr& scale, bool adaptive_sample,
                         bool load_sdf, bool save_sdf) {
                 if (R.size() != 9) throw std::invalid_argument("R is 3 x 3");
                 Matrix3 Rm;
                 for (int r = 0; r < 3; ++r)
                     for (int c = 0; c < 3; ++c) Rm(r, c) = R.data()[3 * r + c];
                 BodyMeshObj::TransformType tt;
                 if (transform_type == "BODY_TO_JOINT") tt = BodyMeshObj::BODY_TO_JOINT;
                 else if (transform_type == "OBJ_TO_JOINT") tt = BodyMeshObj::OBJ_TO_JOINT;
                 else throw std::invalid_argument("transform_type: BODY_TO_JOINT or OBJ_TO_JOINT without a joint");
                 return new BodySDFObj(nullptr, nullptr, filename, Rm, vec3(p), dx, res, col_th, tt, density, vec3(scale),
                                       adaptive_sample, load_sdf, save_sdf);
             }),
             py::arg("filename"), py::arg("R"), py::arg("p"), py::arg("dx"), py::arg("res"), py::arg("col_th"),
             py::arg("transform_type") = "BODY_TO_JOINT", py::arg("density") = 1.0,
             py::arg("scale") = std::vector<double>{1.0, 1.0, 1.0}, py::arg("adaptive_sample") = false,
             py::arg("load_sdf") = false, py::arg("save_sdf") = false,
             "BodySDFObj(sim, joint, filename, R, p, dx, res, col_th, transform_type, density, scale, "
             "adaptive_sample, load_sdf, save_sdf) for a body without a joint")
        .def("save_sdf", &BodySDFObj::save_SDF_file)
        .def("clear_saved_SDF", &BodySDFObj::clear_saved_SDF)
        .def("surface_contact_pt", [](BodySDFObj& b, const arr& xw) { return fixed_out(b.surface_contact_pt(vec3(xw))); })
        .def("test_collision_derivatives", &BodySDFObj::test_collision_derivatives)
        .def("test_collision_derivatives_runtime", [](BodySDFObj& b, const arr& xw, const arr& xw_dot) {
                 return b.test_collision_derivatives_runtime(vec3(xw), vec3(xw_dot));
             })
        .def("get_vertices", [](const BodySDFObj& b) { return points_out(b._V, true); })
        .def("get_faces", [](const BodySDFObj& b) { return faces_out(b._F); })
        .def_property_readonly("mass", [](const BodySDFObj& b) { return b._mass; })
        .def_property_readonly("inertia", [](const BodySDFObj& b) { return fixed_out(b._Inertia); })
        .def_property_readonly("E_ji", [](const BodySDFObj& b) { return fixed_out(b._E_ji); })
        .def_property_readonly("E_oi", [](const BodySDFObj& b) { return fixed_out(b._E_oi); })
        .def("get_AABB", [](const BodySDFObj& b) {
                 auto box = b.get_AABB();
                 return py::make_tuple(fixed_out(box.first), fixed_out(box.second));
             })
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

    m.def("load_obj", [](const std::string& filename, const arr& scale) {
              TriMesh t = load_obj(filename, vec3(scale));
              return py::make_tuple(points_out(t.V, true), faces_out(t.F));
          }, py::arg("filename"), py::arg("scale") = std::vector<double>{1.0, 1.0, 1.0},
          "BodyMeshObj::load_mesh: (vertices (n, 3), faces (3, m))");
    m.def("process_mesh", [](const arr& V, const py::array_t<int64_t, py::array::c_style | py::array::forcecast>& F,
                             double density) {
              TriMesh t;
              t.V = points_of(V);
              if (F.ndim() != 2 || F.shape(1) != 3) throw std::invalid_argument("faces must have shape (m, 3)");
              for (py::ssize_t i = 0; i < F.shape(0); ++i)
                  t.F.push_back({(int)F.at(i, 0), (int)F.at(i, 1), (int)F.at(i, 2)});
              dtype vol;
              Vector3 com;
              Matrix3 I;
              compute_mass_property(t, vol, com, I);
              MassFrame mf = process_mesh(t, density);
              py::dict d;
              d["volume"] = vol;
              d["com"] = fixed_out(com);
              d["I_unit_mass"] = fixed_out(I);
              d["mass"] = mf.mass;
              d["inertia"] = fixed_out(mf.inertia);
              d["E_oi"] = fixed_out(mf.E_oi);
              d["V_body"] = points_out(t.V, true);
              return d;
          }, py::arg("V"), py::arg("F"), py::arg("density") = 1.0,
          "BodyMeshObj::compute_mass_property + process_mesh (volume, centre of mass, principal frame)");
    m.def("so3_derivatives", [](const arr& r) {
              if (r.size() != 3) throw std::invalid_argument("r has 3 components");
              So3Derivs d;
              so3_derivatives(r.data(), d);
              py::dict o;
              o["R"] = row_major(d.R, {3, 3});
              o["Jr"] = row_major(d.Jr, {3, 3});
              o["dR"] = row_major(&d.dR[0][0], {3, 3, 3});
              o["dJr"] = row_major(&d.dJr[0][0], {3, 3, 3});
              o["d2R"] = row_major(&d.d2R[0][0][0], {3, 3, 3, 3});
              o["d2Jr"] = row_major(&d.d2Jr[0][0][0], {3, 3, 3, 3});
              return o;
          }, py::arg("r"), "exp([r]x), the right Jacobian and their first and second derivatives in r");

    py::class_<Simulation>(m, "Simulation",
                           "Scenes of mesh / SDF bodies on fixed, translational or free3d-exp joints with "
                           "general_SDF_contact forces, stepped by BDF1 (redmax_py.Simulation subset)")
        .def(py::init([](const std::string& xml, const std::string& asset_folder, bool verbose, bool precise) {
                 return new Simulation(xml, asset_folder, verbose, precise);
             }),
             py::arg("xml_string"), py::arg("asset_folder"), py::arg("verbose") = false, py::arg("precise") = false)
        .def(py::init([](const std::string& path, bool verbose) { return Simulation::from_file(path, verbose).release(); }),
             py::arg("xml_file_path"), py::arg("verbose") = false)
        .def_readonly("ndof_r", &Simulation::_ndof_r)
        .def_readonly("ndof_m", &Simulation::_ndof_m)
        .def_property_readonly("timestep", [](const Simulation& s) { return s._options->_h; })
        .def_property_readonly("gravity", [](const Simulation& s) { return fixed_out(s._options->_gravity); })
        .def_property_readonly("integrator", [](const Simulation& s) { return s._options->_integrator; })
        .def("init", &Simulation::init)
        .def("get_q", [](const Simulation& s) { return vec_out(s.get_q()); })
        .def("get_qdot", [](const Simulation& s) { return vec_out(s.get_qdot()); })
        .def("get_q_init", [](const Simulation& s) { return vec_out(s.get_q_init()); })
        .def("get_qdot_init", [](const Simulation& s) { return vec_out(s.get_qdot_init()); })
        .def("set_state_init", [](Simulation& s, const arr& q, const arr& qd) { s.set_state_init(vec_in(q), vec_in(qd)); },
             py::arg("q_init"), py::arg("qdot_init"))
        .def("set_q_init", [](Simulation& s, const arr& q) { s.set_q_init(vec_in(q)); }, py::arg("q_init"))
        .def("set_qdot_init", [](Simulation& s, const arr& q) { s.set_qdot_init(vec_in(q)); }, py::arg("qdot_init"))
        .def("set_state", [](Simulation& s, const arr& q, const arr& qd) { s.set_state(vec_in(q), vec_in(qd)); s.update_robot(); },
             py::arg("q"), py::arg("qdot"))
        .def("set_q", [](Simulation& s, const arr& q) { s.set_q(vec_in(q)); s.update_robot(); }, py::arg("q"))
        .def("set_qdot", [](Simulation& s, const arr& q) { s.set_qdot(vec_in(q)); s.update_robot(); }, py::arg("qdot"))
        .def("update_robot", &Simulation::update_robot, py::arg("design_gradient") = false)
        .def("get_q_his", [](const Simulation& s) {
                 std::vector<py::array_t<double>> out;
                 for (const auto& q : s.get_q_his()) out.push_back(vec_out(q));
                 return out;
             })
        .def("get_qdot_his", [](const Simulation& s) {
                 std::vector<py::array_t<double>> out;
                 for (const auto& q : s.get_qdot_his()) out.push_back(vec_out(q));
                 return out;
             })
        .def("get_joint_q", [](const Simulation& s, const std::string& n) { return vec_out(s.get_joint_q(n)); }, py::arg("name"))
        .def("get_joint_qdot", [](const Simulation& s, const std::string& n) { return vec_out(s.get_joint_qdot(n)); },
             py::arg("name"))
        .def("set_joint_q", [](Simulation& s, const std::string& n, const arr& q) { s.set_joint_q(n, vec_in(q)); },
             py::arg("name"), py::arg("q"))
        .def("zero_joint_q", &Simulation::zero_joint_q, py::arg("name"))
        .def("zero_joint_qdot", &Simulation::zero_joint_qdot, py::arg("name"))
        .def("get_body_mass", &Simulation::get_body_mass, py::arg("name"))
        .def("get_body_E0i", [](const Simulation& s, const std::string& n) { return fixed_out(s.get_body_E0i(n)); },
             py::arg("name"))
        .def("get_body_Ei0", [](const Simulation& s, const std::string& n) { return fixed_out(s.get_body_Ei0(n)); },
             py::arg("name"))
        .def("get_body_vertices", [](const Simulation& s, const std::string& n, bool world) {
                 return points_out(s.get_body_vertices(n, world), false);
             }, py::arg("name"), py::arg("world_frame") = false)
        .def("get_body_faces", [](const Simulation& s, const std::string& n) { return faces_out(s.get_body_faces(n)); },
             py::arg("name"))
        .def("get_body_contact_points", [](const Simulation& s, const std::string& n) {
                 return points_out(s.get_body_contact_points(n), true);
             }, py::arg("name"))
        .def("get_body_inertia", [](const Simulation& s, const std::string& n) {
                 for (const auto& b : s._bodies)
                     if (b.name == n) return fixed_out(b.body->_Inertia);
                 throw std::invalid_argument("no body named " + n);
             }, py::arg("name"))
        .def("get_body_Eji", [](const Simulation& s, const std::string& n) {
                 for (const auto& b : s._bodies)
                     if (b.name == n) return fixed_out(b.body->_E_ji);
                 throw std::invalid_argument("no body named " + n);
             }, py::arg("name"))
        .def("get_body_sdf", [](const Simulation& s, const std::string& n) {
                 for (const auto& b : s._bodies)
                     if (b.name == n) {
                         if (!b.sdf) throw std::invalid_argument(n + " is not an SDF body");
                         const BodySDFObj& g = *b.sdf;
                         py::array_t<float> v(std::vector<py::ssize_t>{g._SDF.nk, g._SDF.nj, g._SDF.ni});
                         float* w = v.mutable_data();
                         for (size_t e = 0; e < g._SDF.a.size(); ++e) w[e] = (float)g._SDF.a[e];
                         py::dict d;
                         d["values"] = v;
                         d["min_box"] = fixed_out(g._min_box);
                         d["max_box"] = fixed_out(g._max_box);
                         d["dxyz"] = std::vector<double>{g._dx, g._dy, g._dz};
                         d["col_th"] = g._col_th;
                         return d;
                     }
                 throw std::invalid_argument("no body named " + n);
             }, py::arg("name"), "the body's SDF grid (values (nk, nj, ni) as the GPU holds them, box, spacing, col_th)")
        .def("set_body_external_force", [](Simulation& s, const std::string& n, const arr& f) {
                 s.set_body_external_force(n, v6(f));
             }, py::arg("name"), py::arg("force"))
        .def("get_body_distance", &Simulation::get_body_distance, py::arg("name_from"), py::arg("name_to"))
        .def("body_in_contact", &Simulation::body_in_contact, py::arg("name_a"), py::arg("name_b"), py::arg("eps") = 1e-5)
        .def("get_contact_bodies", &Simulation::get_contact_bodies, py::arg("name"))
        .def("clear_contact_bodies", &Simulation::clear_contact_bodies, py::arg("name"))
        .def("clear_all_saved_sdfs", &Simulation::clear_all_saved_sdfs)
        .def("set_contact_scale", &Simulation::set_contact_scale, py::arg("scale"))
        .def("set_broad_phase", &Simulation::set_broad_phase, py::arg("on"))
        .def("reset", &Simulation::reset, py::arg("backward_flag") = false, py::arg("backward_design_params_flag") = true)
        .def("forward", &Simulation::forward, py::arg("num_steps"), py::arg("verbose") = false,
             py::arg("test_derivatives") = false, py::call_guard<py::gil_scoped_release>())
        .def("is_converged", &Simulation::is_converged)
        .def("body_names", [](const Simulation& s) {
                 std::vector<std::string> n;
                 for (const auto& b : s._bodies) n.push_back(b.name);
                 return n;
             })
        .def("forward_candidates", [](Simulation& s, const arr& q, const arr& qdot, const arr& forces, int num_steps) {
                 const py::ssize_t K = q.ndim() == 2 ? q.shape(0) : -1;
                 const py::ssize_t nb = (py::ssize_t)s._bodies.size();
                 if (K < 0 || q.shape(1) != s._ndof_r || qdot.ndim() != 2 || qdot.shape(0) != K ||
                     qdot.shape(1) != s._ndof_r)
                     throw std::invalid_argument("q and qdot must have shape (candidates, ndof_r)");
                 if (forces.ndim() != 3 || forces.shape(0) != K || forces.shape(1) != nb || forces.shape(2) != 6)
                     throw std::invalid_argument("forces must have shape (candidates, bodies, 6)");
                 std::vector<VectorX> qs(K), qds(K);
                 std::vector<std::vector<Vector6>> fs(K, std::vector<Vector6>(nb));
                 for (py::ssize_t c = 0; c < K; ++c) {
                     qs[c] = VectorX(s._ndof_r);
                     qds[c] = VectorX(s._ndof_r);
                     for (int i = 0; i < s._ndof_r; ++i) {
                         qs[c](i) = q.data()[c * s._ndof_r + i];
                         qds[c](i) = qdot.data()[c * s._ndof_r + i];
                     }
                     for (py::ssize_t b = 0; b < nb; ++b)
                         for (int k = 0; k < 6; ++k) fs[c][b](k) = forces.data()[(c * nb + b) * 6 + k];
                 }
                 std::vector<int> conv;
                 {
                     py::gil_scoped_release rel;
                     s.forward_candidates(qs, qds, fs, num_steps, &conv);
                 }
                 py::array_t<double> qo({K, (py::ssize_t)s._ndof_r}), qdo({K, (py::ssize_t)s._ndof_r});
                 for (py::ssize_t c = 0; c < K; ++c)
                     for (int i = 0; i < s._ndof_r; ++i) {
                         qo.mutable_data()[c * s._ndof_r + i] = qs[c](i);
                         qdo.mutable_data()[c * s._ndof_r + i] = qds[c](i);
                     }
                 py::dict d;
                 d["q"] = qo;
                 d["qdot"] = qdo;
                 d["converged"] = conv;
                 return d;
             }, py::arg("q"), py::arg("qdot"), py::arg("forces"), py::arg("num_steps"),
             "candidates of this scene (own state and external forces per body) stepped together")
        .def("computeMatrices", [](Simulation& s, bool derivs) {
                 py::dict d;
                 MatrixX M, K, D;
                 VectorX f;
                 if (derivs) {
                     std::vector<MatrixX> dM;
                     s.computeMatrices(M, f, dM, K, D);
                     std::vector<py::array_t<double>> dMs;
                     for (const auto& x : dM) dMs.push_back(mat_out(x));
                     d["K"] = mat_out(K);
                     d["D"] = mat_out(D);
                     d["dM_dq"] = dMs;
                 } else {
                     s.computeMatrices(M, f);
                 }
                 d["M"] = mat_out(M);
                 d["f"] = vec_out(f);
                 return d;
             }, py::arg("derivatives") = false, "M, f (and K = df/dq, D = df/dqdot, dM/dq) at the current state")
        .def("last_contact_eval", [](const Simulation& s) {
                 const auto& e = s.last_contact_eval();
                 const py::ssize_t n = (py::ssize_t)e.pairs.size();
                 std::vector<std::pair<std::string, std::string>> names;
                 for (const auto& pr : e.pairs) names.push_back({s._bodies[pr.first].name, s._bodies[pr.second].name});
                 py::array_t<double> f({n, (py::ssize_t)12});
                 std::copy(e.f.begin(), e.f.begin() + 12 * n, f.mutable_data());
                 py::array_t<int32_t> na(n);
                 std::copy(e.na.begin(), e.na.begin() + n, na.mutable_data());
                 py::dict d;
                 d["pairs"] = names;
                 d["f"] = f;
                 d["n_active"] = na;
                 d["derivatives"] = e.derivs;
                 if (e.derivs) {
                     py::array_t<double> K({n, (py::ssize_t)12, (py::ssize_t)12}), D({n, (py::ssize_t)12, (py::ssize_t)12});
                     std::copy(e.K.begin(), e.K.begin() + 144 * n, K.mutable_data());
                     std::copy(e.D.begin(), e.D.begin() + 144 * n, D.mutable_data());
                     d["K"] = K;
                     d["D"] = D;
                 }
                 return d;
             }, "the last contact evaluation: (general body, SDF body) names, wrench f [tau1 f1 tau2 f2] "
                "per pair, active counts and, with derivatives, K = df/d[q1 q2] and D = df/d[phi1 phi2]")
        .def("contact_vjp", [](Simulation& s, const arr& g) {
                 const py::ssize_t n = (py::ssize_t)s.last_contact_eval().pairs.size();
                 if (g.ndim() != 2 || g.shape(0) != n || g.shape(1) != 12)
                     throw std::invalid_argument("g must have shape (pairs of the last evaluation, 12)");
                 std::vector<double> gv(g.data(), g.data() + 12 * n), r;
                 {
                     py::gil_scoped_release rel;
                     r = s.contact_vjp(gv);
                 }
                 py::dict d;
                 const char* keys[4] = {"q1", "q2", "phi1", "phi2"};
                 for (int k = 0; k < 4; ++k) {
                     py::array_t<double> a({n, (py::ssize_t)6});
                     for (py::ssize_t i = 0; i < n; ++i)
                         for (int c = 0; c < 6; ++c) a.mutable_data()[i * 6 + c] = r[26 * i + 6 * k + c];
                     d[keys[k]] = a;
                 }
                 py::array_t<double> kn(n), kt(n);
                 for (py::ssize_t i = 0; i < n; ++i) {
                     kn.mutable_data()[i] = r[26 * i + 24];
                     kt.mutable_data()[i] = r[26 * i + 25];
                 }
                 d["kn"] = kn;
                 d["kt"] = kt;
                 return d;
             }, py::arg("g"), "the adjoint of the last contact evaluation for d(loss)/df per pair: d(loss)/d q1, q2, "
                "phi1, phi2 (6 each), kn, kt")
        .def("stats", [](const Simulation& s) {
                 const auto& st = s.stats();
                 py::dict d;
                 d["steps"] = st.steps;
                 d["evaluations"] = st.evaluations;
                 d["derivative_evaluations"] = st.derivative_evaluations;
                 d["newton_iterations"] = st.newton_iterations;
                 d["pairs_evaluated"] = st.pairs_evaluated;
                 d["pairs_culled"] = st.pairs_culled;
                 d["batches_built"] = st.batches_built;
                 d["eval_ms"] = st.eval_ms;
                 d["step_ms"] = st.step_ms;
                 return d;
             })
        .def("reset_stats", &Simulation::reset_stats);
}
