// The fields of a rigid body that the SDF-contact force reads and writes, with the
// reference's names and meaning (simulation/redmax/Body/Body.h:23-129):
// _E_0i (pose in the world), _phi (body-frame twist [w; v]), _index (maximal-
// coordinate indices of its 6 DoF), _contact_scale, _contact_points and the
// contact-body history.
#pragma once
#include <set>
#include <string>
#include <vector>
#include "Common.h"

namespace redmax {

class Simulation;

class Body {
public:
    Simulation* _sim = nullptr;
    std::string _name;
    std::vector<Vector3> _contact_points;
    std::set<std::string> _contact_body_history;
    std::vector<int> _index;   // 6 consecutive indices into fm / Km / Dm
    dtype _contact_scale = 1.0;
    SE3 _E_0i = SE3::Identity();
    se3 _phi;
    int _p6_column = -1;       // column of the contact-scale design parameter in dfm_dp (-1: inactive)

    Body(Simulation* sim, std::string name, std::vector<Vector3> contact_points, int index0)
        : _sim(sim), _name(std::move(name)), _contact_points(std::move(contact_points)) {
        for (int k = 0; k < 6; ++k) _index.push_back(index0 + k);
    }
    virtual ~Body() = default;

    std::vector<Vector3> get_contact_points() const { return _contact_points; }
    void add_contact_body(std::string body_name) { _contact_body_history.insert(body_name); }
    std::vector<std::string> get_contact_bodies() const {
        return std::vector<std::string>(_contact_body_history.begin(), _contact_body_history.end());
    }
    void clear_contact_bodies() { _contact_body_history.clear(); }
};

}  // namespace redmax
