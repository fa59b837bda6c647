// Host-side shape tables and the polyhedron constructor (surfaces, edges,
// Gauss map) that the reference runs when a convex primitive is created.
#pragma once
#include <stdint.h>

#include <utility>
#include <vector>

namespace cpx {

struct HullData {
	std::vector<double> pts;  // body frame, xyz
	std::vector<std::vector<int>> faces;
	std::vector<int> sid, ncalc, ndir;
	std::vector<double> gnorm;
	std::vector<int> red, minpid, maxpid, refidx;
	std::vector<double> gmin, gmax, grefval;
	std::vector<std::pair<int, int>> arcs, edges;
};

// Static triangle mesh as StaticMeshGeometry(const MeshInput&) builds it
// (Geometry.cpp:678-820), in the mesh body's frame.  Per triangle: normal,
// vertices, edge outward directions, Gauss region (up to 3 directions),
// edge ids, concave neighbour count, restitution / kinetic / static friction.
struct MeshTriangle {
	double n[3], v[9], eout[9], gauss[9], mat[3];
	int eid[3], ngauss, concave;
};
struct HostMesh {
	std::vector<MeshTriangle> tris;
};
HostMesh build_mesh(int n_pts, const double* pts, int n_tris, const int64_t* tri_idx, const double* tri_mat3);

// World copy of a mesh body (StaticMeshGeometry::recomputeFromReference,
// Geometry.cpp:849-857) with its AABB tree built as AABBTree<unsigned int>
// builds it: leaves added in triangle order, surface-area cost, no margin
// (AABB_Tree.h:39-103, :357-424).  Device record layouts: triangle MT_REC
// doubles (normal, vertices, edge out directions, Gauss region, aabb, mat)
// and MT_IREC ints (edge ids, Gauss region size, concave count, local id);
// node box[6] and (left, right), a leaf being (-1, triangle).
#define MT_REC 40
#define MT_IREC 8
struct WorldMesh {
	std::vector<double> tri;   // [T][MT_REC]
	std::vector<int> tri_i;    // [T][MT_IREC]
	std::vector<double> box;   // [N][6]
	std::vector<int> lr;       // [N][2]
	int root = -1, depth = 0;
	double aabb[6];            // StaticMeshGeometry::genAABB (Geometry.cpp:859-868)
};
WorldMesh place_mesh(const HostMesh& m, const double* com3, const double* q4);

struct HostShapes {
	std::vector<int> type, pt_off{0}, face_off{0}, arc_off{0}, edge_off{0};
	std::vector<double> ref_pts;
	std::vector<int> fv_off{0}, fv, ncalc, ndir, sid;
	std::vector<double> gnorm;
	std::vector<int> red, minpid, maxpid, refidx;
	std::vector<double> gmin, gmax, grefval;
	std::vector<int> arcs, edges;
	std::vector<int> vaf_off{0}, vaf, vae_off{0}, vae;
	std::vector<double> radius, height;

	int add_hull(const HullData& d);
	int add_sphere(const double* center, double r);
	// Cylinder as the reference stores it: params8 = centre xyz, height axis
	// xyz, radius, height; gauss42 = 14 Gauss vertices; top24 = top 8-gon.
	// Stored as points [centre, top 8-gon, bottom 8-gon] and normals [axis,
	// Gauss vertices] (the bottom polygon is derived per step).
	int add_cylinder(const double* params8, const double* gauss42, const double* top24, const double* bot24);
	// Static mesh: no convex geometry; the triangles live in meshes[mesh_of[shape]].
	std::vector<int> mesh_of;
	std::vector<HostMesh> meshes;
	int add_mesh(HostMesh m);
};

// Polyhedron(points, surface_vertex_indices) (ConvexPrimitive.cpp:582-658)
// with Surface construction (:783-858) and computeGaussMap (:861-896).
// Cylinder(center, radius, height, axis (0,1,0), 12, 8) (ConvexPrimitive.cpp:222-256)
// in the layout add_cylinder takes.
void build_cylinder(const double* center, double radius, double height, double* params8, double* gauss42, double* top24, double* bot24);

HullData build_polyhedron(const std::vector<double>& pts, const std::vector<std::vector<int>>& faces);

// Mass, inertia tensor about the centre of mass and the centre of mass of a
// closed polyhedron of uniform density as calculateMassProperties computes
// them (projection integrals, RigidBody.cpp:427-640).
void polyhedron_mass_properties(const std::vector<double>& pts, const std::vector<std::vector<int>>& faces, double density,
                                double* mass, double* tensor9, double* com3);
// Mat3::inverse (Mat3.h:57-68).
void mat3_inverse(const double* m9, double* out9);
// The polyhedron of a body created from Polyhedron(points, faces): Gauss map
// of the points as given, body-frame points shifted to the centre of mass;
// mp19 = mass, reference tensor, its inverse.
HullData polyhedron_body(const std::vector<double>& pts, const std::vector<std::vector<int>>& faces, double density, double* mp19);

}  // namespace cpx
