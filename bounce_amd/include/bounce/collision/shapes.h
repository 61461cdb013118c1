// bounce/collision/shapes.h -- hull geometry and collision shapes of the drop-in API.
//
// Class names and public members follow the reference (include/bounce/collision/geometry/hull.h,
// box_hull.h; include/bounce/collision/shapes/{shape,sphere_shape,capsule_shape,hull_shape}.h):
// b3Hull points at caller-owned arrays and is borrowed by b3HullShape ("must remain in scope while
// the shape is in use"); b3BoxHull owns the storage of a box.
#pragma once
#include "../common/math.h"

struct b3Face {
    uint32 edge;
};

struct b3HalfEdge {
    uint32 origin;
    uint32 twin;
    uint32 face;
    uint32 prev;
    uint32 next;
};

struct b3Plane {
    b3Vec3 normal;
    scalar offset;
};

// A convex polyhedron: vertices, half-edges (twins at 2k, 2k + 1), faces and face planes.
struct b3Hull {
    b3Vec3 centroid;
    uint32 vertexCount;
    b3Vec3* vertices;
    uint32 edgeCount;
    b3HalfEdge* edges;
    uint32 faceCount;
    b3Face* faces;
    b3Plane* planes;

    const b3Vec3& GetVertex(uint32 i) const { return vertices[i]; }
    const b3HalfEdge* GetEdge(uint32 i) const { return edges + i; }
    const b3Face* GetFace(uint32 i) const { return faces + i; }
    const b3Plane& GetPlane(uint32 i) const { return planes[i]; }
};

// A box with half-extents (ex, ey, ez) centred at the origin; the same topology as the reference's
// b3BoxHull, so feature ids agree.
struct b3BoxHull : public b3Hull {
    b3Vec3 boxVertices[8];
    b3HalfEdge boxEdges[24];
    b3Face boxFaces[6];
    b3Plane boxPlanes[6];

    b3BoxHull() {}
    b3BoxHull(scalar ex, scalar ey, scalar ez) { SetExtents(ex, ey, ez); }
    void SetIdentity();  // half-extents 1
    void SetExtents(scalar ex, scalar ey, scalar ez);
};

class b3Shape {
public:
    enum Type { e_sphere = 0, e_capsule = 1, e_triangle = 2, e_hull = 3, e_mesh = 4, e_maxShapes = 5 };
    virtual ~b3Shape() {}
    Type GetType() const { return m_type; }
    Type m_type;
    scalar m_radius;
protected:
    explicit b3Shape(Type t, scalar r) : m_type(t), m_radius(r) {}
};

class b3SphereShape : public b3Shape {
public:
    b3SphereShape() : b3Shape(e_sphere, scalar(0)) { m_center.SetZero(); }
    b3Vec3 m_center;
};

class b3CapsuleShape : public b3Shape {
public:
    b3CapsuleShape() : b3Shape(e_capsule, scalar(0)) { m_vertex1.SetZero(); m_vertex2.SetZero(); }
    b3Vec3 m_vertex1, m_vertex2;
};

class b3HullShape : public b3Shape {
public:
    b3HullShape() : b3Shape(e_hull, scalar(0)), m_hull(nullptr) {}  // hulls have no skin
    const b3Hull* m_hull;
};
