#!/bin/sh
# Regenerates the golden fixtures from the compiled reference engine.
# Runs only where the reference checkout exists (REF, default /root/reference);
# the GPU box uses the committed fixtures.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
make -C "$ROOT/oracle" ref REF="${REF:-/root/reference}"
H="$ROOT/oracle/_ref/cp_ref_harness"
cd "$HERE"
rm -rf pile3 c0_full mix3 cyl3 cyl4 chain3 chain3h0 terrain hulls sleep joints joints_h0 limits limits_h0 powered powered_h0 springs springs_h0
# C0 pattern, 3x3x3 cubes: full per-step dumps (state, caches, manifolds, islands).
"$H" pile --n 3 --steps 40 --record 0,1,2,10,39 --out pile3
# C0 in full (10x10x10 cubes): first step and a settled step (no mid/post caches).
"$H" pile --n 10 --steps 121 --record 0,120 --lite 1 --out c0_full
# Mixed pile (cubes and spheres alternating, 3x3x3): sphere-sphere, sphere-box, box-box, ground pairs.
"$H" pile --mixed 1 --n 3 --steps 60 --record 0,5,30,59 --out mix3
# Boxes, spheres and tilted cylinders (3x3x3 and 4x3x4): poly-cylinder (both
# orders), sphere-cylinder, cylinder-cylinder and cylinder-ground pairs.
"$H" pile --mixed 2 --n 3 --steps 80 --record 0,10,30,60,79 --out cyl3
"$H" pile --mixed 2 --n 4 --ny 3 --steps 150 --record 40,90,149 --seed 23 --out cyl4
# Jointed chains (3 chains x 6 cylinder links, ball-sockets with every third
# joint a hinge) falling onto the ground box; with and without the holonomic
# block solver.
"$H" chain --chains 3 --links 6 --steps 100 --record 1,2,45,70,99 --seed 4242 --out chain3
"$H" chain --chains 3 --links 6 --steps 100 --record 1,45,70,99 --seed 4242 --hol 0 --out chain3h0
# Heightfield terrain (12 x 12 cells, 288 triangles, value-noise heights)
# under a box/sphere/cylinder pile and two jointed chains: convex-triangle SAT,
# several manifolds per body pair, the mesh AABB tree's candidate order.
"$H" terrain --grid 12 --cell 0.5 --amp 0.3 --n 2 --chains 2 --links 4 --steps 240 --record 1,60,120,180,239 --rays 96 --out terrain
# Irregular convex hulls (8-32 points on ellipsoids, the mixed-primitive
# library's generator) built by the reference from points and faces, with
# spheres and cylinders, falling onto the ground box: shape construction
# (Gauss maps, mass properties), irregular hull-hull SAT and clipping; rays
# cast at the final state (also on the terrain fixture).
python3 make_hull_input.py hull_input.txt 24 20240
"$H" hulls --in hull_input.txt --n 4 --ny 3 --steps 240 --record 0,60,100,149,239 --rays 128 --out hulls
# Sleeping on (the engine's default parameters): a cube pile, a lone cube and
# a sliding cube fall asleep; a sphere dropped on a sleeping column of three
# cubes wakes it; a waking action on a pile cube before step 250 wakes the
# pile.  Every step's sleep state is recorded.
"$H" sleep --steps 420 --record 0,100,135,136,153,154,200,220,221,250,251,300,313,314,315,330,339,419 --out sleep
# The other equality joints: a chain hanging from a fixed anchor through a
# distance joint, a weld, a slider and a sliding hinge (one holonomic system
# of degrees 1, 6, 5, 4), a welded pair and a slider pair on the ground; with
# and without the holonomic block solver.
"$H" joints --steps 120 --record 0,1,2,30,60,119 --out joints
"$H" joints --steps 120 --record 0,1,30,60,119 --hol 0 --out joints_h0
# Cone and twist limits: a two-link ball-socket pendulum under cone limits
# (0.15 and 0.12 rad) and a hinged block under a twist limit (-0.3, 0.4),
# kicked so that the limits engage and release (steps 15, 16, 32 and 57 solve
# a limit that the step's own update found inside its range); with and
# without the holonomic block solver.
"$H" limits --steps 150 --record 0,1,10,11,14,15,16,20,21,30,32,40,57,80,145,149 --out limits
"$H" limits --steps 150 --record 0,1,10,15,16,20,30,32,57,80,149 --hol 0 --out limits_h0
# Motors and a piston: a motor on a hinge (angle limits -1.0, 1.2) through
# constant torque, target velocity, target position and off; a motor on a
# sliding hinge at a target velocity; a piston on a slider (slide limits
# -0.4, 0.9) through target velocity, target position, constant force and
# off; the mode changes are dumped as drive_events.
"$H" powered --steps 200 --record 0,1,27,29,32,40,54,70,80,89,93,120,130,141,160,199 --out powered
"$H" powered --steps 200 --record 0,27,40,54,70,80,93,130,160,199 --hol 0 --out powered_h0
# Springs (a box hanging from a fixed anchor, a second box hanging from it,
# two boxes on the ground pushed apart) and joints with a custom CFM (the
# middle joint of a ball-socket chain in a holonomic system, a hinge).
"$H" springs --steps 200 --record 0,1,10,50,99,100,150,199 --out springs
"$H" springs --steps 200 --record 0,1,50,100,199 --hol 0 --out springs_h0
