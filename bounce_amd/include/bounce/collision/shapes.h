// bounce/collision/shapes.h -- hull geometry and collision shapes of the drop-in API.
//
// Class names and public members follow the reference (include/bounce/collision/geometry/hull.h,
// box_hull.h, mesh.h; include/bounce/collision/shapes/{shape,sphere_shape,capsule_shape,
// triangle_shape,hull_shape,mesh_shape}.h): b3Hull and b3Mesh point at caller-owned arrays and are
// borrowed by b3HullShape and b3MeshShape ("must remain in scope while the shape is in use");
// b3BoxHull owns the storage of a box.
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

// A triangle with the hull skin radius (triangle_shape.h). Its edge wings are not passed to the
// device: a standalone triangle keeps m_hasE*Vertex false, as b3TriangleShape::Set leaves it.
class b3TriangleShape : public b3Shape {
public:
    b3TriangleShape() : b3Shape(e_triangle, scalar(0)) {
        m_vertex1.SetZero();
        m_vertex2.SetZero();
        m_vertex3.SetZero();
        m_hasE1Vertex = m_hasE2Vertex = m_hasE3Vertex = false;
    }
    void Set(const b3Vec3& v1, const b3Vec3& v2, const b3Vec3& v3) {
        m_vertex1 = v1;
        m_vertex2 = v2;
        m_vertex3 = v3;
        m_hasE1Vertex = m_hasE2Vertex = m_hasE3Vertex = false;
    }
    b3Vec3 m_vertex1, m_vertex2, m_vertex3;
    bool m_hasE1Vertex, m_hasE2Vertex, m_hasE3Vertex;
    b3Vec3 m_e1Vertex, m_e2Vertex, m_e3Vertex;
};

// Mesh triangle (mesh.h): wing vertices u1..u3 and vertices v1..v3.
struct b3MeshTriangle {
    uint32& GetWingVertex(uint32 i) { return (&u1)[i]; }
    uint32 GetWingVertex(uint32 i) const { return (&u1)[i]; }
    uint32& GetVertex(uint32 i) { return (&v1)[i]; }
    uint32 GetVertex(uint32 i) const { return (&v1)[i]; }
    uint32 u1, u2, u3;
    uint32 v1, v2, v3;
};

// A triangle mesh over caller-owned arrays. The device library builds the mesh's AABB tree and
// edge wings when a world using it is uploaded, as b3Mesh::BuildTree and BuildAdjacency do, so the
// two calls here only exist for source compatibility.
struct b3Mesh {
    uint32 vertexCount;
    b3Vec3* vertices;
    uint32 triangleCount;
    b3MeshTriangle* triangles;
    void BuildTree() {}
    void BuildAdjacency() {}
    const b3Vec3& GetVertex(uint32 index) const { return vertices[index]; }
    const b3MeshTriangle* GetTriangle(uint32 index) const { return triangles + index; }
};

// A scaled triangle mesh with the hull skin radius (mesh_shape.h).
class b3MeshShape : public b3Shape {
public:
    b3MeshShape() : b3Shape(e_mesh, scalar(0)), m_mesh(nullptr) { m_scale.Set(scalar(1), scalar(1), scalar(1)); }
    const b3Mesh* m_mesh;
    b3Vec3 m_scale;
};
