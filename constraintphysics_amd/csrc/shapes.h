// Host-side shape tables and the polyhedron constructor (surfaces, edges,
// Gauss map) that the reference runs when a convex primitive is created.
#pragma once
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
};

// Polyhedron(points, surface_vertex_indices) (ConvexPrimitive.cpp:582-658)
// with Surface construction (:783-858) and computeGaussMap (:861-896).
// Cylinder(center, radius, height, axis (0,1,0), 12, 8) (ConvexPrimitive.cpp:222-256)
// in the layout add_cylinder takes.
void build_cylinder(const double* center, double radius, double height, double* params8, double* gauss42, double* top24, double* bot24);

HullData build_polyhedron(const std::vector<double>& pts, const std::vector<std::vector<int>>& faces);

// Mass, inertia tensor about the centre of mass and the centre of mass of a
// closed polyhedron of uniform density (divergence theorem over the
// tetrahedra spanned by the origin and a fan triangulation of each face).
void polyhedron_mass_properties(const std::vector<double>& pts, const std::vector<std::vector<int>>& faces, double density,
                                double* mass, double* tensor9, double* com3);

}  // namespace cpx
