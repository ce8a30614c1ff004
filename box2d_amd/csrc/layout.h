// Data layout shared by the host bookkeeping and the gfx950 kernels.
//
// Everything the step reads or writes on the GPU lives in arrays indexed by stable ids (body id,
// shape id, contact id, joint id), so a body falling asleep or a contact moving between the
// non-touching set and a graph color never moves its data: only the small index lists that say
// which ids take part in a stage change. Static bodies keep an identity body state, which is the
// "dummy state" the reference substitutes for a null solver index (src/body.h:165), so solver
// lanes never branch on null bodies.
//
// Contact constraints are stored structure-of-arrays (one float array per field) so a wavefront
// reads each field with one coalesced load; the fields are those of b2ContactConstraintWide
// (reference src/contact_solver.c:1070-1100).
#pragma once

#include "box2d/collision.h"
#include "box2d/math_functions.h"
#include "box2d/types.h"

#include <stdint.h>

namespace b2g
{

// Body flags (same bit values as reference src/body.h:11-57).
enum BodyFlags : uint32_t
{
	kLockLinearX = 0x00000001,
	kLockLinearY = 0x00000002,
	kLockAngularZ = 0x00000004,
	kIsFast = 0x00000008,
	kIsBullet = 0x00000010,
	kIsSpeedCapped = 0x00000020,
	kHadTimeOfImpact = 0x00000040,
	kAllowFastRotation = 0x00000080,
	kEnlargeBounds = 0x00000100,
	kDynamicFlag = 0x00000200,
	kDirtyMass = 0x00000400,
	kEnableSleep = 0x00000800,
	kEnableContactRecycling = 0x00001000,
	kAllLocks = kLockAngularZ | kLockLinearX | kLockLinearY,
	kTransientFlags = kIsFast | kIsSpeedCapped | kHadTimeOfImpact,
};

// Contact flags (same bit values as reference src/contact.h:14-42).
enum ContactFlags : uint32_t
{
	kContactTouching = 0x00000001,
	kContactHitEvent = 0x00000002,
	kContactEnableContactEvents = 0x00000004,
	kContactRecycle = 0x00000008,
	kSimTouching = 0x00010000,
	kSimDisjoint = 0x00020000,
	kSimStartedTouching = 0x00040000,
	kSimStoppedTouching = 0x00080000,
	kSimEnableHitEvent = 0x00100000,
	kSimEnablePreSolveEvents = 0x00200000,
	kSimRelativeTransformValid = 0x00400000,
	// the contact's friction and restitution come from user mixing callbacks and are out of date:
	// its next full narrowphase update asks the host for new values
	kSimMixStale = 0x00800000,
};

// Shape flags
enum ShapeFlags : uint32_t
{
	kShapeEnableHitEvents = 0x1,
	kShapeEnableContactEvents = 0x2,
	kShapeEnablePreSolveEvents = 0x4,
	kShapeIsSensor = 0x8,
	kShapeEnableSensorEvents = 0x10,
	kShapeHasProxy = 0x20, // the shape is in a broadphase tree (its body is not disabled)
};

struct Softness
{
	float biasRate;
	float massScale;
	float impulseScale;
};

// Per body (indexed by body id). Mirrors the solver-relevant part of b2BodySim.
struct BodySim
{
	b2Transform transform;
	b2Vec2 center;
	b2Rot rotation0;
	b2Vec2 center0;
	b2Vec2 localCenter;
	b2Vec2 force;
	float torque;
	float invMass;
	float invInertia;
	float minExtent;
	float maxExtent;
	float linearDamping;
	float angularDamping;
	float gravityScale;
	uint32_t flags;
	int headShapeId;
};

// Per body solver state, 32 bytes (reference src/body.h:152-171).
struct alignas( 16 ) BodyState
{
	b2Vec2 linearVelocity;
	float angularVelocity;
	uint32_t flags;
	b2Vec2 deltaPosition;
	b2Rot deltaRotation;
};

// Per body data the finalize stage updates and the host reads back for sleep and broadphase.
struct BodyAux
{
	float sleepThreshold;
	float sleepTime;
	float safetyFactor;
	uint32_t flags; // b2Body flags including the per-step transient bits
	int type;
	int enlarged; // set by finalize when any shape AABB escaped its fat AABB
	int pad0;
	int pad1;
};

// Per shape (indexed by shape id).
struct ShapeDev
{
	int type;
	int bodyId;
	int nextShapeId;
	uint32_t flags;
	float friction;
	float restitution;
	float rollingResistance;
	float tangentSpeed;
	float aabbMargin;
	int enlargedAABB;
	b2Vec2 localCentroid;
	b2AABB aabb;
	b2AABB fatAABB;
	union
	{
		b2Capsule capsule;
		b2Circle circle;
		b2Polygon polygon;
		b2Segment segment;
		b2ChainSegment chainSegment;
	};
	b2Filter filter;
};

// Per contact (indexed by contact id). Mirrors b2ContactSim (reference src/contact.h:103-142)
// without the transient solver indices, which the GPU derives from body ids.
struct ContactSim
{
	int shapeIdA;
	int shapeIdB;
	int bodyIdA;
	int bodyIdB;
	uint32_t simFlags;
	// decisions of the last narrowphase update, read by step recording: bit 0 manifold recycled;
	// bits 2-3 and 4-5 warm-start source of point 0 and 1 (0 none, 1 + old point index)
	uint32_t stepInfo;
	b2Rot cachedRotationA;
	b2Rot cachedRotationB;
	b2Transform cachedRelativePose;
	float friction;
	float restitution;
	float rollingResistance;
	float tangentSpeed;
	b2SimplexCache cache;
	b2Manifold manifold;
};

// A contact created on the host this step, scattered into the device contact pool.
struct ContactInit
{
	int contactId;
	int shapeIdA;
	int shapeIdB;
	int bodyIdA;
	int bodyIdB;
	uint32_t simFlags;
	float friction;
	float restitution;
};

// Revolute joint plus the shared joint header (reference src/joint.h:267-301 and the
// b2RevoluteJoint payload at src/joint.h:196-222).
struct JointSim
{
	int jointId;
	int bodyIdA;
	int bodyIdB;
	int type;
	b2Transform localFrameA;
	b2Transform localFrameB;
	float invMassA, invMassB;
	float invIA, invIB;
	float constraintHertz;
	float constraintDampingRatio;
	Softness constraintSoftness;
	float forceThreshold;
	float torqueThreshold;

	b2Vec2 linearImpulse;
	float springImpulse;
	float motorImpulse;
	float lowerImpulse;
	float upperImpulse;
	float hertz;
	float dampingRatio;
	float targetAngle;
	float maxMotorTorque;
	float motorSpeed;
	float lowerAngle;
	float upperAngle;
	int indexA; // body id whose state the solver uses (static bodies read their identity state)
	int indexB;
	b2Transform frameA;
	b2Transform frameB;
	b2Vec2 deltaCenter;
	float axialMass;
	Softness springSoftness;
	int enableSpring;
	int enableMotor;
	int enableLimit;
	int slotA; // LDS slots of the two bodies in the fused solver
	int slotB;

	// Fields of the other joint types (src/joint.h:64-200). The revolute fields above are shared
	// where a type has the same quantity: linearImpulse (weld), springImpulse, motorImpulse,
	// lowerImpulse, upperImpulse, hertz, dampingRatio, motorSpeed, springSoftness (prismatic,
	// wheel, distance), maxMotorTorque (wheel), axialMass (weld, prismatic, wheel, distance).
	float lowerTranslation; // prismatic, wheel
	float upperTranslation;
	float targetTranslation; // prismatic
	float maxMotorForce;	 // prismatic, distance
	b2Vec2 pointImpulse;	 // prismatic: perpendicular and angular impulse
	float perpImpulse;		 // wheel
	float perpMass;
	float motorMass;

	float length; // distance
	float minLength;
	float maxLength;
	float lowerSpringForce;
	float upperSpringForce;
	float distanceImpulse;
	b2Vec2 anchorA;
	b2Vec2 anchorB;

	float linearHertz; // weld, motor
	float linearDampingRatio;
	float angularHertz;
	float angularDampingRatio;
	Softness linearSpring;
	Softness angularSpring;
	float angularImpulse; // weld

	b2Vec2 linearVelocity; // motor
	float maxVelocityForce;
	float angularVelocity;
	float maxVelocityTorque;
	float maxSpringForce;
	float maxSpringTorque;
	b2Vec2 linearVelocityImpulse;
	float angularVelocityImpulse;
	b2Vec2 linearSpringImpulse;
	float angularSpringImpulse;
	b2Mat22 linearMass;
	float angularMass;

	b2Vec2 maxVelocityForce2; // mover: per-axis force limits
	float pointMass;		  // mover and pogo: scalar linear effective mass

	b2Vec2 normal; // pogo
	float restLength;
	float maxTensionForce;
	float maxCompressionForce;
	float pogoImpulse;
	float pogoVelocity;
	int eventFlag; // a biased solve of this step reached the force or torque threshold
	int pad;
};

// Contact constraint fields, one device array per field (structure of arrays).
enum ConstraintField
{
	kcInvMassA,
	kcInvMassB,
	kcInvIA,
	kcInvIB,
	kcNormalX,
	kcNormalY,
	kcFriction,
	kcTangentSpeed,
	kcRollingResistance,
	kcRollingMass,
	kcRollingImpulse,
	kcBiasRate,
	kcMassScale,
	kcImpulseScale,
	kcRestitution,
	kcAnchorA1X,
	kcAnchorA1Y,
	kcAnchorB1X,
	kcAnchorB1Y,
	kcBaseSeparation1,
	kcNormalImpulse1,
	kcTangentImpulse1,
	kcTotalNormalImpulse1,
	kcNormalMass1,
	kcTangentMass1,
	kcRelativeVelocity1,
	kcAnchorA2X,
	kcAnchorA2Y,
	kcAnchorB2X,
	kcAnchorB2Y,
	kcBaseSeparation2,
	kcNormalImpulse2,
	kcTangentImpulse2,
	kcTotalNormalImpulse2,
	kcNormalMass2,
	kcTangentMass2,
	kcRelativeVelocity2,
	kcFieldCount
};

// Step parameters shared by all stages of one step (reference b2StepContext, src/solver.h:120-185).
struct StepParams
{
	float dt;
	float inv_dt;
	float h;
	float inv_h;
	int subStepCount;
	Softness contactSoftness;
	Softness staticSoftness;
	float restitutionThreshold;
	float maxLinearVelocity;
	float contactSpeed;
	float contactHertz;
	float contactDampingRatio;
	float hitEventThreshold;
	b2Vec2 gravity;
	float contactRecycleDistance;
	int enableWarmStarting;
	int enableSleep;
	int enableContinuous;
	int enableContactSoftening;
	// user mixing callbacks: bit 0 friction, bit 1 restitution (the host evaluates them; the
	// narrowphase keeps the contact's values instead of mixing the shapes' own)
	int customMixing;
	// some joint has a finite force or torque threshold: the pack kernel gathers joint events
	int jointEvents;
	// the world has a pre-solve callback: the narrowphase leaves contacts that enable pre-solve
	// events to the host
	int preSolve;
	// the world has a sensor on a static body: continuous collision records its hits
	int sensorHits;
	// continuous collision runs on the host for every fast body (a pre-continuous callback with
	// shapes that enable pre-solve events): the device kernel is not launched
	int hostContinuous;
};

} // namespace b2g
