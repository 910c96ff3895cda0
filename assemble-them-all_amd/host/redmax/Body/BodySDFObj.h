// SDF body of the standalone host layer: the reference's public grid fields
// (simulation/redmax/Body/BodySDFObj.h:13-16: _min_box, _max_box, _dx/_dy/_dz,
// _col_th, _SDF in sdfgen's Array3 layout a[i + ni*(j + nj*k)]) plus the query
// methods distance / collision (BodySDFObj.h:35-76), evaluated on the GPU.
// The grid is given directly, read from the reference's .sdf text format
// (BodySDFObj.cpp:86-116), or built from a triangle mesh on the GPU
// (precompute_SDF, BodySDFObj.cpp:38-77).
#pragma once
#include <array>
#include <string>
#include <vector>
#include "Body/BodyMeshObj.h"
#include "atac.h"

namespace sdfgen {
template <class T>
struct Array3 {
    int ni = 0, nj = 0, nk = 0;
    std::vector<T> a;
    void resize(int i, int j, int k) {
        ni = i; nj = j; nk = k;
        a.assign((size_t)i * j * k, T(0));
    }
    T& operator()(int i, int j, int k) { return a[(size_t)i + (size_t)ni * ((size_t)j + (size_t)nj * (size_t)k)]; }
    const T& operator()(int i, int j, int k) const { return a[(size_t)i + (size_t)ni * ((size_t)j + (size_t)nj * (size_t)k)]; }
};
}  // namespace sdfgen

namespace redmax {

class BodySDFObj : public BodyMeshObj {
public:
    Vector3 _min_box, _max_box;      // bounding box of the SDF grid
    dtype _dx = 0, _dy = 0, _dz = 0; // grid spacing
    dtype _col_th = 0;               // collision threshold
    sdfgen::Array3<dtype> _SDF;      // voxel grid of the SDF

    // BodySDFObj(sim, joint, filename, R, p, dx, res, col_th, transform_type, density, scale,
    // adaptive_sample, load_sdf, save_sdf) (BodySDFObj.h:18-27, BodySDFObj.cpp:6-36): the mesh
    // body, then its grid from <filename without .obj>.sdf when load_sdf finds it, else built
    // by precompute_SDF (and written there when save_sdf)
    BodySDFObj(Simulation* sim, Joint* joint, std::string filename, Matrix3 R, Vector3 p, dtype dx, int res,
               dtype col_th, TransformType transform_type = BODY_TO_JOINT, dtype density = 1.0,
               Vector3 scale = Vector3::Ones(), bool adaptive_sample = false, bool load_sdf = false,
               bool save_sdf = false);
    // a body given by its grid and contact points directly
    BodySDFObj(Simulation* sim, std::string name, const sdfgen::Array3<dtype>& sdf, Vector3 min_box, Vector3 max_box,
               dtype dx, dtype dy, dtype dz, dtype col_th, std::vector<Vector3> contact_points, int index0);
    ~BodySDFObj() override;

    // the .sdf cache next to the mesh file (BodySDFObj.cpp:79-121)
    bool load_SDF();
    void save_SDF() const;
    void clear_saved_SDF() const;
    std::string sdf_filename() const;

    // reads / writes the reference's .sdf text format (ni nj nk / min_box / max_box /
    // dx dy dz / values in a[i + ni*(j + nj*k)] order); false when the file is unreadable
    static bool load_SDF_file(const std::string& path, sdfgen::Array3<dtype>& sdf, Vector3& min_box, Vector3& max_box,
                              dtype& dx, dtype& dy, dtype& dz);
    void save_SDF_file(const std::string& path) const;

    // the grid precompute_SDF builds from a closed triangle mesh (BodySDFObj.cpp:38-77):
    // spacing min(dx, bounding-box extent / res) per axis, one cell of padding around
    // the FP32 bounds, values from the level-set construction (sdfgen::make_level_set3)
    // run on the GPU by atac_sdf_from_mesh; throws on a bad mesh
    static void precompute_SDF(const std::vector<Vector3>& V, const std::vector<std::array<int, 3>>& F, dtype dx,
                               int res, sdfgen::Array3<dtype>& sdf, Vector3& min_box, Vector3& max_box, dtype& dx_out,
                               dtype& dy_out, dtype& dz_out);

    // query_SDF(R2^T (xw - p2)) + col_th, FLT_MAX + col_th outside the box (BodySDFObj.cpp:207-213)
    dtype distance(Vector3 xw);
    // x - d e in the body frame, e the unit SDF gradient (zero where |grad| < 1e-5)
    // (BodySDFObj.cpp:215-237)
    Vector3 surface_contact_pt(Vector3 xw);
    void collision(Vector3 xw, Vector3 xw_dot, dtype& d, Vector3& n, dtype& ddot, Vector3& tdot, Vector3& xi2);
    void collision(Vector3 xw, Vector3 xw_dot, dtype& d, Vector3& n, dtype& ddot, Vector3& tdot, Vector3& xi2,
                   RowVector3& dd_dxw, RowVector6& dd_dq2, Matrix3& dn_dxw, Matrix36& dn_dq2, RowVector3& dddot_dxw,
                   RowVector3& dddot_dxwdot, RowVector6& dddot_dq2, RowVector6& dddot_dphi2, Matrix3& dtdot_dxw,
                   Matrix3& dtdot_dxwdot, Matrix36& dtdot_dq2, Matrix36& dtdot_dphi2, Matrix3& dxi2_dxw,
                   Matrix36& dxi2_dq2);

    // the reference's self-tests of the collision derivatives (BodySDFObj.cpp:385-737):
    // every derivative block against central finite differences of the collision values,
    // printed per block; they return the largest relative error. The first one draws a
    // random pose, twist and a point inside the body from a fixed seed.
    dtype test_collision_derivatives();
    dtype test_collision_derivatives_runtime(Vector3 xw, Vector3 xw_dot);

    // FP32 copy of the grid in HBM, created on first use; call after changing _SDF
    atac_grid device_grid();
    void release_device_grid();

private:
    atac_grid grid_ = nullptr;
    void collision_values(const Vector3& xw, const Vector3& xw_dot, double* values, double* derivs);
};

}  // namespace redmax
