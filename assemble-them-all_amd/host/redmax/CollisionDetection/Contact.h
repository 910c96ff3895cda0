// Contact record (simulation/redmax/CollisionDetection/Contact.h:6-20).
#pragma once
#include "Common.h"

namespace redmax {

class Contact {
public:
    Vector3 _xi;      // contact point in the contact body's frame
    Vector3 _xw;      // contact point in the world frame
    dtype _d;         // penetration depth
    Vector3 _normal;  // contact normal
    int _id;          // index of the contact point in the body

    Contact(Vector3 xi, Vector3 xw, dtype d, Vector3 normal, int id = 0)
        : _xi(xi), _xw(xw), _d(d), _normal(normal), _id(id) {}
};

}  // namespace redmax
