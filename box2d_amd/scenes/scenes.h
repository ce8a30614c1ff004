// Scene driver for the Box2D step path.
//
// Written only against the public Box2D C API (box2d/box2d.h), so the same source compiles
// against the reference library and against box2d_amd, and drives both through identical
// calls. Each scene is addressed by a small integer handle so that callers (Python ctypes,
// C benchmarks) never need Box2D's struct layouts.
#pragma once

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined( _WIN32 )
#define B2S_API __declspec( dllexport )
#else
#define B2S_API __attribute__( ( visibility( "default" ) ) )
#endif

// Scene kinds
enum
{
	// Triangular box pyramid on a static ground box.
	// iparams: [0] rows, [1] enableSleep. fparams: [0] ground half width (half height is 1).
	// Follows the reference benchmark construction (shared/benchmarks.c:88-131).
	B2S_PYRAMID = 0,

	// Falling hinged boxes used by the reference determinism test (shared/determinism.c:11-107).
	B2S_FALLING_HINGES = 1,

	// A single dynamic box resting on a ground box (test/test_world.c:19-98 "HelloWorld").
	B2S_HELLO = 2,

	// Small mixed scene: circles, capsules, polygons and rounded boxes dropped on a ground box.
	// iparams: [0] body count, [1] enableSleep, [2] seed. fparams: unused.
	B2S_MIXED_PILE = 3,

	// Mixed-shape rain into a static U-shaped chain container: equal quarters of circles
	// (r in [0.2, 0.5]), capsules (half-length in [0.2, 0.6], r in [0.1, 0.25]), convex hulls of
	// 3-8 points in a 1 m disc, and rounded boxes (radius 0.05), densities in [0.5, 2], on a grid
	// of `columns` columns spanning a column of the given width (spacing width / columns, rows
	// stacked upward from 110 m). The container is an open chain whose interior is on the
	// segments' right side (walls 100 m high, floor 2 m segments).
	// iparams: [0] body count, [1] columns, [2] enableSleep, [3] seed. fparams: [0] column width.
	B2S_RAIN = 4,

	// One world of the batched uneven-worlds workload, sized to about a target body count, with
	// config-0 physics (gravity -10, friction 0.6, sleep off, ground box 40 x 1 half extents).
	// iparams: [0] target bodies, [1] kind (0 box pyramid, 1 random box/circle pile, 2 hanging
	// chains of capsule links on revolute joints), [2] seed.
	// Pyramids take the largest row count whose box count fits the target; piles place the
	// target count on a grid; chains are 1-8 chains of 5-20 links (count = chains * links).
	B2S_BATCH_WORLD = 5,

	// Differentiable push: a pyramid of 1 m boxes (iparams[0] rows, default 10: 55 boxes) on the
	// 40 x 1 m ground box, sleep off, and a 1 m pusher box on the ground 0.5 m left of the pyramid
	// with initial velocity (fparams[0], fparams[1]), spin fparams[2] and friction fparams[3].
	// Bodies: the pyramid row by row (top box last), then the pusher.
	B2S_PUSH = 6,

	// Every joint type on a 80 x 1 m ground box (half extents), with the type's spring, limit
	// and motor settings varied so each solver branch runs, and bodies that also touch the ground
	// or each other: weld cantilevers (rigid and soft), prismatic sliders, wheel cars driving on
	// the ground, distance ropes (rigid, spring, limits, motor), motor-joint followers, filter
	// joints between overlapping boxes, mover and pogo joints, and revolute hinges with spring,
	// limit and motor.
	// iparams: [0] enableSleep, [1] event thresholds (optional): the joints in ascending id get
	// user data k + 1 and, by k mod 4, thresholds that report every awake step, a force of 40 N,
	// a torque of 2 N m, or never; before step 151 every joint moves one place in that cycle.
	// Bodies are listed in creation order.
	B2S_JOINT_ZOO = 7,

	// Contact events: circles, boxes and capsules thrown onto a 40 x 1 m ground box (half extents)
	// from a grid of heights with random velocities. Contact events are on for two shapes in
	// three and for the ground, hit events for every other shape. Sleep is on.
	// iparams: [0] body count, [1] seed.
	B2S_EVENTS = 8,

	// Edits between steps: 24 boxes, circles and capsules on a 30 x 1 m ground box (half extents),
	// a five-link hinge chain and a kinematic paddle. Before chosen steps b2s_step moves bodies,
	// applies forces and impulses, replaces shape geometry and recomputes mass, changes the
	// hinges' motor, spring, limit and frame, destroys a joint and two bodies, creates bodies,
	// and drives the paddle with b2Body_SetTargetTransform; after each step it folds getter
	// results (local/world points and vectors, joint angles, motor torques, joint and shape
	// queries, masses) into b2s_probe_hash. iparams: [0] enableSleep.
	B2S_EDITS = 9,

	// Setters and callbacks between steps: 30 boxes, circles and capsules (user data = index,
	// user material ids 0-7, some with custom filtering, one with a second shape), a three-link
	// hinge chain and a kinematic paddle on a 30 x 1 m ground box (half extents). The world has
	// friction and restitution callbacks and a custom filter; before chosen steps b2s_step changes
	// damping, gravity scale, shape materials, density and filters, body types (dynamic, static,
	// kinematic, with joints attached), disables and enables bodies, destroys a shape, changes
	// sleep settings and wakes or sleeps bodies, swaps and clears the callbacks and changes the
	// restitution and hit-event thresholds. After each step it folds getter results (damping,
	// gravity scale, materials, filters, types, sleep settings, thresholds, contact data of a body,
	// a shape and a contact) into b2s_probe_hash. iparams: [0] enableSleep.
	B2S_SETTERS = 10,

	// Pre-solve: a static one-way platform (16 x 0.5 m at y = 4, pre-solve events on) over a
	// 30 x 1 m ground box (half extents); 12 boxes launched up through it from below and 12
	// boxes, circles and capsules dropped onto it, every third with pre-solve events. The
	// callback drops the points of platform contacts approached from below and the second point
	// of some other pairs. Before chosen steps b2s_step turns the platform's flag off and on,
	// relaunches bodies and clears and restores the callback. After each step it folds the
	// callback's call count and an order-independent sum of its arguments, the platform's flag
	// and its contact count into b2s_probe_hash. iparams: [0] enableSleep.
	B2S_PRESOLVE = 11,

	// Sensors: a static zone (mask excluding category 4) and a thin static strip on a 40 x 1 m
	// ground box (half extents) with a wall at x = 32, a kinematic sensor circle sweeping across,
	// a dynamic carrier with a sensor circle, 30 visitors dropped over the zone and the carrier
	// (two in three sense, every fifth in category 4, every seventh custom-filtered: odd user data
	// is not sensed) and three projectiles crossing the strip within one step. Before chosen steps
	// b2s_step toggles the zone's sensor events, destroys a visitor, disables and enables the
	// carrier, destroys the carried sensor, turns a visitor's events off, changes a filter and
	// fires the projectiles again. After each step it folds the sensor events and the sensors'
	// overlaps into b2s_probe_hash. Listed bodies: the kinematic sensor, the carrier, the 30
	// visitors, the projectiles. iparams: [0] enableSleep.
	B2S_SENSORS = 12,

	// Bullets: five bullets (circle, box, capsule, a circle with sensor and pre-solve events, a
	// hexagon) and two fast bodies that are not bullets, fired at 200-260 m/s through a stack of
	// twelve boxes, a thin standing plank, a rotating kinematic paddle, a dynamic body carrying a
	// sensor ring and a static ghost wall with pre-solve events, inside a 70 x 12 m box of static
	// walls. Before chosen steps b2s_step fires them again from either end and at angles, turns a
	// fast body into a bullet and a bullet off and on again. After each step it folds the sensor
	// events, the ring's overlaps, the bullet flags and the pre-continuous callback's tally into
	// b2s_probe_hash. Listed bodies: the bullets, the two fast bodies, the stack, the plank, the
	// paddle and the ring's body. iparams: [0] enableSleep, [1] pre-continuous callback (hits on
	// the ghost wall above y = 3 are ignored).
	B2S_BULLETS = 13,
};

// Create a world with workerCount threads (0 or 1 means serial) and build a scene in it.
// Returns a handle >= 0, or -1 on failure.
B2S_API int b2s_create( int kind, int workerCount, const int* iparams, int iparamCount, const float* fparams, int fparamCount );

// The b2WorldId behind a handle packed as index1 | generation << 16 (0 if the handle is invalid).
B2S_API uint32_t b2s_world_id( int handle );

// Destroy the world behind a handle.
B2S_API void b2s_destroy( int handle );

// Step the world count times with the given time step and sub-step count.
B2S_API void b2s_step( int handle, float timeStep, int subStepCount, int count );

// Wall-clock milliseconds spent in count steps (monotonic clock around b2World_Step calls).
B2S_API double b2s_step_timed( int handle, float timeStep, int subStepCount, int count );

// b2s_step_timed that also adds each step's b2World_GetProfile fields (23 floats, b2Profile
// order) into profileSums[0..22] after the step; the profile read sits outside the clock.
B2S_API double b2s_step_profiled( int handle, float timeStep, int subStepCount, int count, double* profileSums );

// Body ids (index1 - 1 of each b2BodyId) of the scene bodies in creation order. Returns the count.
B2S_API int b2s_body_ids( int handle, int* out, int capacity );

// Number of bodies created by the scene (static ground excluded).
B2S_API int b2s_body_count( int handle );

// djb2 hash (b2Hash) of the world transforms of the scene bodies in creation order,
// the same construction as the reference determinism check (shared/determinism.c:110-125).
B2S_API uint32_t b2s_hash( int handle );

// Getter results folded after every step of a B2S_EDITS scene (FNV-1a), 0 for other scenes.
B2S_API uint32_t b2s_probe_hash( int handle );

// b2World_GetStateHash of the scene's world, 0 for an unknown handle.
B2S_API uint64_t b2s_state_hash( int handle );

// Copy the world transforms (p.x, p.y, q.c, q.s) of the scene bodies in creation order.
// Returns the number of bodies written.
B2S_API int b2s_get_transforms( int handle, float* out, int capacity );

// Copy the linear and angular velocities (v.x, v.y, w) of the scene bodies.
B2S_API int b2s_get_velocities( int handle, float* out, int capacity );

// Body move event count and awake body count of the last step.
B2S_API int b2s_move_count( int handle );
B2S_API int b2s_awake_body_count( int handle );

// The contact events of the last step, flattened in 10-word records: begin (kind 0), end (kind 1)
// and hit (kind 2) events in the order b2World_GetContactEvents lists them. A record is
// [kind, shapeA index1, shapeB index1, contact index1, contact generation, point x, point y,
// normal x, normal y, approach speed] with floats as their bits (zero for begin and end events).
// Writes at most capacity / 10 records and returns the record count.
B2S_API int b2s_contact_events( int handle, uint32_t* out, int capacity );

// Joint events of the last step in the order b2World_GetJointEvents lists them: records of
// [joint index1, joint generation, user data as an integer]. Writes at most capacity / 3 records
// and returns the record count.
B2S_API int b2s_joint_events( int handle, uint32_t* out, int capacity );

// Sensor events of the last step, begin events then end events as b2World_GetSensorEvents lists
// them: records of [0 begin / 1 end, sensor index1, sensor generation, visitor index1, visitor
// generation]. Writes at most capacity / 5 records and returns the record count.
B2S_API int b2s_sensor_events( int handle, uint32_t* out, int capacity );

// World counters: out[0..5] = bodyCount, shapeCount, contactCount, jointCount, islandCount,
// awakeContactCount; out[6..29] = colorCounts[0..23]. Returns the number of ints written (30).
B2S_API int b2s_counters( int handle, int* out, int capacity );

// Run the reference determinism protocol on the falling hinges scene: step at 1/60 with 4
// sub-steps until every body sleeps or stepLimit is reached.
// Writes the sleep step and the transform hash at that step. Returns 1 if the scene slept.
B2S_API int b2s_falling_hinges( int workerCount, int stepLimit, int* sleepStep, uint32_t* hash );

#ifdef __cplusplus
}
#endif
