// GPU half of b2World_Step for gfx950: device arrays, kernels and the per-step launch sequence.
//
// Stage order per step (reference src/physics_world.c:845-1005, src/solver.c:1010-1198):
//   collide        one lane per awake contact: AABB test, recycling, manifold, warm-start id match
//   prepare        one lane per graph constraint (contacts) / joint
//   sub-steps      integrate velocities; warm start, solve, integrate positions, relax
//                  (overflow constraints first, single lane in order, then one launch per color)
//   restitution    only when some shape has restitution
//   store          impulses back into the contact manifolds
//   finalize       one lane per awake body: transforms, sleep velocity, fat AABBs, enlarge flags
// Colors are Gauss-Seidel phases: constraints inside a color touch disjoint dynamic bodies, so a
// color is solved by one fully parallel launch and the result is independent of lane order.

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "box2d_amd.h"
#include "batch_tuning.h"
#include "device.h"
#include "narrowphase.h"
#include "solver_device.h"
#include "chain.h"
#include "host_threads.h"
#include "toi.h"
#include "world.h"

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

namespace b2g
{

static char s_lastError[256] = "";

const char* LastDeviceError()
{
	return s_lastError;
}

#define HIP_CHECK( expr )                                                                                                        \
	do                                                                                                                           \
	{                                                                                                                            \
		hipError_t _e = ( expr );                                                                                                \
		if ( _e != hipSuccess )                                                                                                  \
		{                                                                                                                        \
			snprintf( s_lastError, sizeof( s_lastError ), "%s failed: %s", #expr, hipGetErrorString( _e ) );                   \
			::b2g::Fatal( "HIP error: %s (%s:%d)", s_lastError, __FILE__, __LINE__ );                                             \
		}                                                                                                                        \
	}                                                                                                                            \
	while ( 0 )

constexpr int kBlock = 256;

static int ColorBlockFromEnv()
{
	const char* v = getenv( "B2AMD_COLOR_BLOCK" );
	int b = v != nullptr ? atoi( v ) : 128;
	return ( b == 64 || b == 128 || b == 256 || b == 512 || b == 1024 ) ? b : 128;
}

static inline int Blocks( int n )
{
	return ( n + kBlock - 1 ) / kBlock;
}

// ---------------------------------------------------------------------------------------------
// Kernels

struct InitArgs
{
	ContactSim* contacts;
	const ContactInit* inits;
	int count;
};

// Scatter newly created contacts into the id-indexed contact array.
B2G_DEV void InitContactsBody( const InitArgs& a, int i )
{
	if ( i >= a.count )
	{
		return;
	}
	ContactSim* contacts = a.contacts;
	ContactInit ci = a.inits[i];
	ContactSim c;
	memset( &c, 0, sizeof( c ) );
	c.shapeIdA = ci.shapeIdA;
	c.shapeIdB = ci.shapeIdB;
	c.bodyIdA = ci.bodyIdA;
	c.bodyIdB = ci.bodyIdB;
	c.simFlags = ci.simFlags;
	c.friction = ci.friction;
	c.restitution = ci.restitution;
	contacts[ci.contactId] = c;
}

struct CollideArgs
{
	const int* list;
	int count;
	ContactSim* contacts;
	const ShapeDev* shapes;
	const BodySim* sims;
	const BodyAux* aux;
	int* changes;
	int* counters; // [0] change count, [1] recycled count, [2] deferred count
	float recycleDistance;
	int customMixing; // StepParams::customMixing
	int preSolve;	  // the world has a pre-solve callback: contacts that enable it go to the host
	int* deferred;	  // their ids
	// the pinned mailbox's change and deferred lists, written beside the device lists (null: the
	// pack copies the device lists)
	int* mailChanges;
	int* mailDeferred;
};

// Where the device narrowphase reports: change records, recycled contacts, and contacts whose
// pre-solve callback must run on the host (they are left untouched and listed).
struct DeviceCollideSink
{
	const CollideArgs& a;

	__device__ void Change( int code ) const
	{
		int slot = atomicAdd( a.counters + 0, 1 );
		a.changes[slot] = code;
		if ( a.mailChanges != nullptr )
		{
			a.mailChanges[slot] = code;
		}
	}
	__device__ void Recycled() const
	{
		atomicAdd( a.counters + 1, 1 );
	}
	__device__ bool CheckRecycle() const
	{
		return true;
	}
	__device__ bool Defer( int contactId, uint32_t simFlags ) const
	{
		if ( a.preSolve != 0 && ( simFlags & kSimEnablePreSolveEvents ) != 0 )
		{
			int slot = atomicAdd( a.counters + 2, 1 );
			a.deferred[slot] = contactId;
			if ( a.mailDeferred != nullptr )
			{
				a.mailDeferred[slot] = contactId;
			}
			return true;
		}
		return false;
	}
	__device__ void PreSolve( const ContactSim&, b2Manifold* ) const
	{
	}
};

B2G_DEV float ShapeRadius( const ShapeDev& s )
{
	switch ( s.type )
	{
		case b2_capsuleShape:
			return s.capsule.radius;
		case b2_circleShape:
			return s.circle.radius;
		case b2_polygonShape:
			return s.polygon.radius;
		default:
			return 0.0f;
	}
}

// Narrowphase update for one contact (reference src/physics_world.c:451-609 and
// src/contact.c:519-698). The sink receives the change records and runs the pre-solve callback;
// the device sink instead defers contacts that need it, before anything is written, and the host
// runs this same function for them with a sink that calls the callback.
template <typename Sink>
B2G_DEV void CollideBody( const CollideArgs& a, int k, const Sink& sink )
{
	if ( k >= a.count )
	{
		return;
	}

	int contactId = a.list[k];
	ContactSim* cs = a.contacts + contactId;
	const ShapeDev& shapeA = a.shapes[cs->shapeIdA];
	const ShapeDev& shapeB = a.shapes[cs->shapeIdB];

	bool overlap = b2AABB_Overlaps( shapeA.fatAABB, shapeB.fatAABB );
	if ( overlap == false )
	{
		cs->simFlags |= kSimDisjoint;
		cs->simFlags &= ~kSimTouching;
		sink.Change( contactId * 4 + 1 );
		return;
	}

	uint32_t simFlags = cs->simFlags;
	bool wasTouching = ( simFlags & kSimTouching ) != 0;

	int bodyIdA = shapeA.bodyId;
	int bodyIdB = shapeB.bodyId;
	const BodySim& simA = a.sims[bodyIdA];
	const BodySim& simB = a.sims[bodyIdB];
	b2Transform transformA = simA.transform;
	b2Transform transformB = simB.transform;
	const BodyAux& auxA = a.aux[bodyIdA];
	const BodyAux& auxB = a.aux[bodyIdB];

	float recycleDistance = a.recycleDistance;
	float recycleDistanceNonTouching = b2MinFloat( recycleDistance, kSpeculativeDistance );

	bool isFast = ( auxA.flags & kIsFast ) || ( auxB.flags & kIsFast );

	if ( sink.CheckRecycle() && isFast == false && recycleDistance > 0.0f && ( simFlags & kSimRelativeTransformValid ) &&
		 ( simFlags & kContactRecycle ) )
	{
		b2Rot cachedQA = cs->cachedRotationA;
		b2Rot cachedQB = cs->cachedRotationB;
		b2Transform xfc = cs->cachedRelativePose;
		b2Transform xf = b2InvMulWorldTransforms( transformA, transformB );

		float cosA = transformA.q.c * cachedQA.c + transformA.q.s * cachedQA.s;
		float cosB = transformB.q.c * cachedQB.c + transformB.q.s * cachedQB.s;
		float minCos = b2MinFloat( cosA, cosB );

		float maxExtentA = auxA.type == b2_staticBody ? 0.0f : simA.maxExtent;
		float maxExtentB = auxB.type == b2_staticBody ? 0.0f : simB.maxExtent;
		float maxExtent = b2MaxFloat( maxExtentA, maxExtentB );
		float distance = b2Distance( xf.p, xfc.p );
		b2Rot qr = b2InvMulRot( xf.q, xfc.q );

		float tolerance = wasTouching ? recycleDistance : recycleDistanceNonTouching;

		if ( minCos > 0.98f && distance + maxExtent * b2AbsFloat( qr.s ) < tolerance )
		{
			b2Rot dqA = b2MulRot( transformA.q, b2InvertRot( cachedQA ) );
			b2Rot dqB = b2MulRot( transformB.q, b2InvertRot( cachedQB ) );
			b2Vec2 normal = cs->manifold.normal;
			b2Vec2 dc = b2SubPos( simB.center, simA.center );

			for ( int i = 0; i < cs->manifold.pointCount; ++i )
			{
				b2ManifoldPoint* mp = cs->manifold.points + i;
				b2Vec2 rA = b2RotateVector( dqA, mp->anchorA );
				b2Vec2 rB = b2RotateVector( dqB, mp->anchorB );
				b2Vec2 dp = b2Add( dc, b2Sub( rB, rA ) );
				mp->separation = mp->baseSeparation + b2Dot( dp, normal );
				mp->persisted = true;
			}
			cs->stepInfo = 1u | ( 1u << 2 ) | ( 2u << 4 );

			sink.Recycled();
			return;
		}
	}

	if ( sink.Defer( contactId, simFlags ) )
	{
		return;
	}

	cs->cachedRotationA = transformA.q;
	cs->cachedRotationB = transformB.q;
	cs->cachedRelativePose = b2InvMulWorldTransforms( transformA, transformB );
	simFlags |= kSimRelativeTransformValid;

	b2Vec2 centerOffsetA = b2RotateVector( transformA.q, simA.localCenter );
	b2Vec2 centerOffsetB = b2RotateVector( transformB.q, simB.localCenter );

	// b2UpdateContact
	b2Manifold oldManifold = cs->manifold;
	b2Transform relativeTransform = b2InvMulWorldTransforms( transformA, transformB );
	b2LocalManifold local = ComputeManifoldCached( shapeA, shapeB, relativeTransform, &cs->cache );

	b2Manifold m;
	memset( &m, 0, sizeof( m ) );
	m.normal = b2RotateVector( transformA.q, local.normal );
	m.pointCount = local.pointCount;

	b2Vec2 originDelta = b2SubPos( transformA.p, transformB.p );
	for ( int i = 0; i < local.pointCount; ++i )
	{
		b2ManifoldPoint* mp = m.points + i;
		mp->anchorA = b2RotateVector( transformA.q, local.points[i].point );
		mp->anchorB = b2Add( mp->anchorA, originDelta );
		mp->separation = local.points[i].separation;
		mp->id = local.points[i].id;
	}

	// mixing at every full update (reference src/contact.c:548-551); user callbacks run on the
	// host, which is asked for new values when the stored ones are out of date (change kind 0)
	if ( ( a.customMixing & 1 ) == 0 )
	{
		cs->friction = sqrtf( shapeA.friction * shapeB.friction );
	}
	if ( ( a.customMixing & 2 ) == 0 )
	{
		cs->restitution = b2MaxFloat( shapeA.restitution, shapeB.restitution );
	}
	if ( a.customMixing != 0 && ( simFlags & kSimMixStale ) != 0 )
	{
		simFlags &= ~kSimMixStale;
		sink.Change( contactId * 4 + 0 );
	}

	if ( shapeA.rollingResistance > 0.0f || shapeB.rollingResistance > 0.0f )
	{
		float radiusA = ShapeRadius( shapeA );
		float radiusB = ShapeRadius( shapeB );
		float maxRadius = b2MaxFloat( radiusA, radiusB );
		cs->rollingResistance = b2MaxFloat( shapeA.rollingResistance, shapeB.rollingResistance ) * maxRadius;
	}
	else
	{
		cs->rollingResistance = 0.0f;
	}

	cs->tangentSpeed = shapeA.tangentSpeed + shapeB.tangentSpeed;

	int pointCount = m.pointCount;
	bool touching = pointCount > 0;

	// pre-solve callback (reference src/contact.c:573-583): it may edit the manifold
	if ( touching && ( simFlags & kSimEnablePreSolveEvents ) != 0 )
	{
		sink.PreSolve( *cs, &m );
		pointCount = m.pointCount;
		touching = pointCount > 0;
	}

	if ( touching && ( ( shapeA.flags | shapeB.flags ) & kShapeEnableHitEvents ) )
	{
		simFlags |= kSimEnableHitEvent;
	}
	else
	{
		simFlags &= ~kSimEnableHitEvent;
	}

	if ( pointCount > 0 )
	{
		m.rollingImpulse = oldManifold.rollingImpulse;
	}

	uint32_t stepInfo = 0;
	for ( int i = 0; i < pointCount; ++i )
	{
		b2ManifoldPoint* mp2 = m.points + i;
		mp2->anchorA = b2Sub( mp2->anchorA, centerOffsetA );
		mp2->anchorB = b2Sub( mp2->anchorB, centerOffsetB );
		mp2->normalImpulse = 0.0f;
		mp2->tangentImpulse = 0.0f;
		mp2->totalNormalImpulse = 0.0f;
		mp2->normalVelocity = 0.0f;
		mp2->persisted = false;

		uint16_t id2 = mp2->id;
		for ( int j = 0; j < oldManifold.pointCount; ++j )
		{
			b2ManifoldPoint* mp1 = oldManifold.points + j;
			if ( mp1->id == id2 )
			{
				mp2->normalImpulse = mp1->normalImpulse;
				mp2->tangentImpulse = mp1->tangentImpulse;
				mp2->persisted = true;
				mp1->normalImpulse = 0.0f;
				mp1->tangentImpulse = 0.0f;
				stepInfo |= (uint32_t)( j + 1 ) << ( 2 + 2 * i );
				break;
			}
		}
	}

	if ( touching )
	{
		simFlags |= kSimTouching;
	}
	else
	{
		simFlags &= ~kSimTouching;
	}

	for ( int i = 0; i < m.pointCount; ++i )
	{
		m.points[i].baseSeparation = m.points[i].separation;
	}

	cs->manifold = m;
	cs->simFlags = simFlags;
	cs->stepInfo = stepInfo;

	if ( touching == true && wasTouching == false )
	{
		sink.Change( contactId * 4 + 2 );
	}
	else if ( touching == false && wasTouching == true )
	{
		sink.Change( contactId * 4 + 3 );
	}
}

struct PrepareArgs
{
	ConstraintSoA c;
	int count;
	const ContactSim* contacts;
	const BodySim* sims;
	const BodyState* states;
	const BodyAux* aux;
	const int* bodySlot;
	StepParams p;
};

// Constraint preparation (reference src/contact_solver.c:1573-1809; the overflow variant at
// :24-160 computes the same values).
B2G_DEV void PrepareContactsBody( const PrepareArgs& a, int i )
{
	if ( i >= a.count )
	{
		return;
	}

	const ConstraintSoA& c = a.c;
	const ContactSim* cs = a.contacts + c.contactId[i];
	const b2Manifold& manifold = cs->manifold;

	int bodyIdA = cs->bodyIdA;
	int bodyIdB = cs->bodyIdB;
	c.bodyA[i] = bodyIdA;
	c.bodyB[i] = bodyIdB;
	c.pointCount[i] = manifold.pointCount;

	bool staticA = a.aux[bodyIdA].type == b2_staticBody;
	bool staticB = a.aux[bodyIdB].type == b2_staticBody;
	c.slotA[i] = staticA ? kStaticSlot : a.bodySlot[bodyIdA];
	c.slotB[i] = staticB ? kStaticSlot : a.bodySlot[bodyIdB];

	float mA = a.sims[bodyIdA].invMass;
	float iA = a.sims[bodyIdA].invInertia;
	float mB = a.sims[bodyIdB].invMass;
	float iB = a.sims[bodyIdB].invInertia;

	b2Vec2 vA = a.states[bodyIdA].linearVelocity;
	float wA = a.states[bodyIdA].angularVelocity;
	b2Vec2 vB = a.states[bodyIdB].linearVelocity;
	float wB = a.states[bodyIdB].angularVelocity;

	c.f[kcInvMassA][i] = mA;
	c.f[kcInvMassB][i] = mB;
	c.f[kcInvIA][i] = iA;
	c.f[kcInvIB][i] = iB;

	{
		float k = iA + iB;
		c.f[kcRollingMass][i] = k > 0.0f ? 1.0f / k : 0.0f;
	}

	Softness soft = a.p.contactSoftness;
	if ( staticA || staticB )
	{
		soft = a.p.staticSoftness;
	}
	else if ( a.p.enableContactSoftening )
	{
		float contactHertz = b2MinFloat( a.p.contactHertz, 0.125f * a.p.inv_h );
		float ratio = 1.0f;
		if ( mA < mB )
		{
			ratio = b2MaxFloat( 0.5f, mA / mB );
		}
		else if ( mB < mA )
		{
			ratio = b2MaxFloat( 0.5f, mB / mA );
		}
		soft = MakeSoft( ratio * contactHertz, ratio * a.p.contactDampingRatio, a.p.h );
	}

	float warmStartScale = a.p.enableWarmStarting ? 1.0f : 0.0f;

	b2Vec2 normal = manifold.normal;
	c.f[kcNormalX][i] = normal.x;
	c.f[kcNormalY][i] = normal.y;
	c.f[kcFriction][i] = cs->friction;
	c.f[kcTangentSpeed][i] = cs->tangentSpeed;
	c.f[kcRestitution][i] = cs->restitution;
	c.f[kcRollingResistance][i] = cs->rollingResistance;
	c.f[kcRollingImpulse][i] = warmStartScale * manifold.rollingImpulse;
	c.f[kcBiasRate][i] = soft.biasRate;
	c.f[kcMassScale][i] = soft.massScale;
	c.f[kcImpulseScale][i] = soft.impulseScale;

	b2Vec2 tangent = b2RightPerp( normal );

	for ( int j = 0; j < 2; ++j )
	{
		int o = j == 0 ? 0 : ( kcAnchorA2X - kcAnchorA1X );
		if ( j < manifold.pointCount )
		{
			const b2ManifoldPoint* mp = manifold.points + j;
			b2Vec2 rA = mp->anchorA;
			b2Vec2 rB = mp->anchorB;
			c.f[kcAnchorA1X + o][i] = rA.x;
			c.f[kcAnchorA1Y + o][i] = rA.y;
			c.f[kcAnchorB1X + o][i] = rB.x;
			c.f[kcAnchorB1Y + o][i] = rB.y;
			c.f[kcBaseSeparation1 + o][i] = mp->separation - b2Dot( b2Sub( rB, rA ), normal );
			c.f[kcNormalImpulse1 + o][i] = warmStartScale * mp->normalImpulse;
			c.f[kcTangentImpulse1 + o][i] = warmStartScale * mp->tangentImpulse;
			c.f[kcTotalNormalImpulse1 + o][i] = 0.0f;

			float rnA = b2Cross( rA, normal );
			float rnB = b2Cross( rB, normal );
			float kNormal = mA + mB + iA * rnA * rnA + iB * rnB * rnB;
			c.f[kcNormalMass1 + o][i] = kNormal > 0.0f ? 1.0f / kNormal : 0.0f;

			float rtA = b2Cross( rA, tangent );
			float rtB = b2Cross( rB, tangent );
			float kTangent = mA + mB + iA * rtA * rtA + iB * rtB * rtB;
			c.f[kcTangentMass1 + o][i] = kTangent > 0.0f ? 1.0f / kTangent : 0.0f;

			b2Vec2 vrA = b2Add( vA, b2CrossSV( wA, rA ) );
			b2Vec2 vrB = b2Add( vB, b2CrossSV( wB, rB ) );
			c.f[kcRelativeVelocity1 + o][i] = b2Dot( normal, b2Sub( vrB, vrA ) );
		}
		else
		{
			c.f[kcBaseSeparation1 + o][i] = 0.0f;
			c.f[kcNormalImpulse1 + o][i] = 0.0f;
			c.f[kcTangentImpulse1 + o][i] = 0.0f;
			c.f[kcTotalNormalImpulse1 + o][i] = 0.0f;
			c.f[kcAnchorA1X + o][i] = 0.0f;
			c.f[kcAnchorA1Y + o][i] = 0.0f;
			c.f[kcAnchorB1X + o][i] = 0.0f;
			c.f[kcAnchorB1Y + o][i] = 0.0f;
			c.f[kcNormalMass1 + o][i] = 0.0f;
			c.f[kcTangentMass1 + o][i] = 0.0f;
			c.f[kcRelativeVelocity1 + o][i] = 0.0f;
		}
	}
}

struct PrepareJointArgs
{
	JointSim* joints;
	const int* list;
	int count;
	const BodySim* sims;
	const BodyAux* aux;
	const int* bodySlot;
	StepParams p;
};

// Joint preparation (reference src/joint.c:1406-1452, src/revolute_joint.c:217-281).
B2G_DEV void PrepareJointsBody( const PrepareJointArgs& a, int k )
{
	if ( k >= a.count )
	{
		return;
	}

	JointSim* joint = a.joints + a.list[k];
	float hertz = b2MinFloat( joint->constraintHertz, 0.25f * a.p.inv_h );
	joint->constraintSoftness = MakeSoft( hertz, joint->constraintDampingRatio, a.p.h );

	const BodySim& simA = a.sims[joint->bodyIdA];
	const BodySim& simB = a.sims[joint->bodyIdB];

	float mA = simA.invMass;
	float iA = simA.invInertia;
	float mB = simB.invMass;
	float iB = simB.invInertia;
	joint->invMassA = mA;
	joint->invMassB = mB;
	joint->invIA = iA;
	joint->invIB = iB;

	joint->indexA = joint->bodyIdA;
	joint->indexB = joint->bodyIdB;
	joint->slotA = a.aux[joint->bodyIdA].type == b2_staticBody ? kStaticSlot : a.bodySlot[joint->bodyIdA];
	joint->slotB = a.aux[joint->bodyIdB].type == b2_staticBody ? kStaticSlot : a.bodySlot[joint->bodyIdB];

	// Type terms (src/revolute_joint.c:217-281, src/weld_joint.c:148-220, src/prismatic_joint.c:295-350,
	// src/wheel_joint.c:197-282, src/distance_joint.c:250-313, src/motor_joint.c:179-245,
	// src/mover_joint.c:47-84, src/pogo_joint.c:98-151).
	int type = joint->type;
	if ( type == b2_filterJoint )
	{
		return;
	}
	if ( type == b2_moverJoint )
	{
		// linear only
		joint->invIA = 0.0f;
		joint->invIB = 0.0f;
		float k = mA + mB;
		joint->pointMass = k > 0.0f ? 1.0f / k : 0.0f;
		if ( a.p.enableWarmStarting == 0 )
		{
			joint->linearVelocityImpulse = { 0.0f, 0.0f };
		}
		return;
	}
	joint->deltaCenter = b2SubPos( simB.center, simA.center );
	if ( type == b2_distanceJoint )
	{
		joint->anchorA = b2RotateVector( simA.transform.q, b2Sub( joint->localFrameA.p, simA.localCenter ) );
		joint->anchorB = b2RotateVector( simB.transform.q, b2Sub( joint->localFrameB.p, simB.localCenter ) );

		b2Vec2 rA = joint->anchorA;
		b2Vec2 rB = joint->anchorB;
		b2Vec2 separation = b2Add( b2Sub( rB, rA ), joint->deltaCenter );
		b2Vec2 axis = b2Normalize( separation );

		float crA = b2Cross( rA, axis );
		float crB = b2Cross( rB, axis );
		float k = mA + mB + iA * crA * crA + iB * crB * crB;
		joint->axialMass = k > 0.0f ? 1.0f / k : 0.0f;
		joint->springSoftness = MakeSoft( joint->hertz, joint->dampingRatio, a.p.h );

		if ( a.p.enableWarmStarting == 0 )
		{
			joint->distanceImpulse = 0.0f;
			joint->lowerImpulse = 0.0f;
			joint->upperImpulse = 0.0f;
			joint->motorImpulse = 0.0f;
		}
		return;
	}

	joint->frameA.q = b2MulRot( simA.transform.q, joint->localFrameA.q );
	joint->frameA.p = b2RotateVector( simA.transform.q, b2Sub( joint->localFrameA.p, simA.localCenter ) );
	joint->frameB.q = b2MulRot( simB.transform.q, joint->localFrameB.q );
	joint->frameB.p = b2RotateVector( simB.transform.q, b2Sub( joint->localFrameB.p, simB.localCenter ) );

	switch ( type )
	{
		case b2_revoluteJoint:
		{
			float k2 = iA + iB;
			joint->axialMass = k2 > 0.0f ? 1.0f / k2 : 0.0f;
			joint->springSoftness = MakeSoft( joint->hertz, joint->dampingRatio, a.p.h );
			if ( a.p.enableWarmStarting == 0 )
			{
				joint->linearImpulse = { 0.0f, 0.0f };
				joint->springImpulse = 0.0f;
				joint->motorImpulse = 0.0f;
				joint->lowerImpulse = 0.0f;
				joint->upperImpulse = 0.0f;
			}
			break;
		}
		case b2_weldJoint:
		{
			float ka = iA + iB;
			joint->axialMass = ka > 0.0f ? 1.0f / ka : 0.0f;
			joint->linearSpring = joint->linearHertz == 0.0f ? joint->constraintSoftness
															 : MakeSoft( joint->linearHertz, joint->linearDampingRatio, a.p.h );
			joint->angularSpring = joint->angularHertz == 0.0f ? joint->constraintSoftness
															   : MakeSoft( joint->angularHertz, joint->angularDampingRatio, a.p.h );
			if ( a.p.enableWarmStarting == 0 )
			{
				joint->linearImpulse = { 0.0f, 0.0f };
				joint->angularImpulse = 0.0f;
			}
			break;
		}
		case b2_prismaticJoint:
		{
			joint->springSoftness = MakeSoft( joint->hertz, joint->dampingRatio, a.p.h );
			if ( a.p.enableWarmStarting == 0 )
			{
				joint->pointImpulse = { 0.0f, 0.0f };
				joint->springImpulse = 0.0f;
				joint->motorImpulse = 0.0f;
				joint->lowerImpulse = 0.0f;
				joint->upperImpulse = 0.0f;
			}
			break;
		}
		case b2_wheelJoint:
		{
			b2Vec2 rA = joint->frameA.p;
			b2Vec2 rB = joint->frameB.p;

			b2Vec2 d = b2Add( joint->deltaCenter, b2Sub( rB, rA ) );
			b2Vec2 axisA = b2RotateVector( joint->frameA.q, b2Vec2{ 1.0f, 0.0f } );
			b2Vec2 perpA = b2LeftPerp( axisA );

			// perpendicular constraint (keeps the wheel on its line)
			float s1 = b2Cross( b2Add( d, rA ), perpA );
			float s2 = b2Cross( rB, perpA );
			float kp = mA + mB + iA * s1 * s1 + iB * s2 * s2;
			joint->perpMass = kp > 0.0f ? 1.0f / kp : 0.0f;

			// spring
			float a1 = b2Cross( b2Add( d, rA ), axisA );
			float a2 = b2Cross( rB, axisA );
			float ka = mA + mB + iA * a1 * a1 + iB * a2 * a2;
			joint->axialMass = ka > 0.0f ? 1.0f / ka : 0.0f;

			joint->springSoftness = MakeSoft( joint->hertz, joint->dampingRatio, a.p.h );

			float km = iA + iB;
			joint->motorMass = km > 0.0f ? 1.0f / km : 0.0f;

			if ( a.p.enableWarmStarting == 0 )
			{
				joint->perpImpulse = 0.0f;
				joint->springImpulse = 0.0f;
				joint->motorImpulse = 0.0f;
				joint->lowerImpulse = 0.0f;
				joint->upperImpulse = 0.0f;
			}
			break;
		}
		case b2_pogoJoint:
		{
			b2Vec2 rA = joint->frameA.p;
			b2Vec2 rB = joint->frameB.p;
			float crA = b2Cross( rA, joint->normal );
			float crB = b2Cross( rB, joint->normal );
			float k = mA + mB + iA * crA * crA + iB * crB * crB;
			joint->pointMass = k > 0.0f ? 1.0f / k : 0.0f;

			// The reference zeroes a spring-less pogo's impulse in its warm start; nothing reads
			// the impulse in between, so it is zeroed here, where the per-body warm start (which
			// only reads joints) does not need to write it.
			if ( a.p.enableWarmStarting == 0 || joint->hertz == 0.0f )
			{
				joint->pogoImpulse = 0.0f;
			}
			break;
		}
		case b2_motorJoint:
		{
			b2Vec2 rA = joint->frameA.p;
			b2Vec2 rB = joint->frameB.p;

			joint->linearSpring = MakeSoft( joint->linearHertz, joint->linearDampingRatio, a.p.h );
			joint->angularSpring = MakeSoft( joint->angularHertz, joint->angularDampingRatio, a.p.h );

			b2Mat22 kl;
			kl.cx.x = mA + mB + rA.y * rA.y * iA + rB.y * rB.y * iB;
			kl.cx.y = -rA.y * rA.x * iA - rB.y * rB.x * iB;
			kl.cy.x = kl.cx.y;
			kl.cy.y = mA + mB + rA.x * rA.x * iA + rB.x * rB.x * iB;
			joint->linearMass = b2GetInverse22( kl );

			float ka = iA + iB;
			joint->angularMass = ka > 0.0f ? 1.0f / ka : 0.0f;

			if ( a.p.enableWarmStarting == 0 )
			{
				joint->linearVelocityImpulse = { 0.0f, 0.0f };
				joint->angularVelocityImpulse = 0.0f;
				joint->linearSpringImpulse = { 0.0f, 0.0f };
				joint->angularSpringImpulse = 0.0f;
			}
			break;
		}
		default:
			break;
	}
}

struct BodyArgs
{
	const int* awake;
	int count;
	BodySim* sims;
	BodyState* states;
	BodyAux* aux;
	StepParams p;
};

__global__ void IntegrateVelocitiesKernel( BodyArgs a )
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if ( i >= a.count )
	{
		return;
	}
	int id = a.awake[i];
	IntegrateVelocity( a.sims[id], a.states + id, a.p );
}

__global__ void IntegratePositionsKernel( BodyArgs a )
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if ( i >= a.count )
	{
		return;
	}
	IntegratePosition( a.states + a.awake[i], a.p );
}

struct WarmArgs
{
	BodyArgs b;
	ConstraintSoA c;
	const JointSim* joints;
	const int* jointList;
	const int* warmStart;
	const int* warmList;
	int overflowContactStart;
};

// Velocity integration and the whole warm-start stage of one substep, one lane per awake body
// (WarmStartBody applies the body's constraint deltas in the stage's order).
// 64-lane workgroups for the per-body substep kernels: a large world's few thousand bodies spread
// over as many compute units as possible (11.6 -> 9.6 us per integrate-and-warm-start launch on
// the 5050-body pyramid, with each lane gathering up to eight constraints' fields before applying
// them)
constexpr int kWarmBlock = 64;

__global__ __launch_bounds__( kWarmBlock ) void IntegrateWarmStartKernel( WarmArgs a )
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if ( i >= a.b.count )
	{
		return;
	}
	int id = a.b.awake[i];
	int begin = a.warmStart[i], end = a.warmStart[i + 1];
	BodyState* state = a.b.states + id;
	// integrate and warm start in registers, one store at the end, so every gather can issue
	// before it
	BodyState st = *state;
	IntegrateVelocity( a.b.sims[id], &st, a.b.p );
	float vx = st.linearVelocity.x, vy = st.linearVelocity.y, w = st.angularVelocity;
	WarmStartBody<8>( a.c, a.joints, a.jointList, a.overflowContactStart, a.warmList, begin, end, GlobalBodies{ a.b.states },
					  st.deltaRotation, st.deltaPosition, vx, vy, w );
	state->linearVelocity = { vx, vy };
	state->angularVelocity = w;
}

enum ColorStage
{
	kStageWarmStart,
	kStageSolve,
	kStageRelax,
	kStageRestitution,
};

struct ColorArgs
{
	ConstraintSoA c;
	JointSim* joints;
	const int* jointList;
	int jointStart;
	int jointCount;
	int contactStart;
	int contactCount;
	BodyState* states;
	StepParams p;
	int stage;
};

// True when any constraint in the 4-lane group of local index k has a nonzero value in field f
// (the wide solver skips or runs rolling and restitution per lane group).
B2G_DEV bool GroupAnyNonZero( const ConstraintSoA& c, int f, int colorStart, int colorCount, int k )
{
	int base = k & ~3;
	int end = b2MinInt( base + 4, colorCount );
	for ( int j = base; j < end; ++j )
	{
		if ( c.f[f][colorStart + j] != 0.0f )
		{
			return true;
		}
	}
	return false;
}

// Lane t of one graph color of one stage: joints occupy the first jointCount lanes, contacts
// the rest (joint and contact blocks of one color touch disjoint bodies).
// Joint and contact index ranges of one color (or of the overflow constraints).
struct ColorRange
{
	int jointStart;
	int jointCount;
	int contactStart;
	int contactCount;
};

template <int Stage, typename Bodies>
B2G_DEV void ColorLane( const ColorArgs& a, const ColorRange& r, int t, const Bodies& B )
{
	if ( t < r.jointCount )
	{
		JointSim* joint = a.joints + a.jointList[r.jointStart + t];
		if ( Stage == kStageWarmStart )
		{
			WarmStartJoint( joint, B );
		}
		else if ( Stage == kStageSolve )
		{
			SolveJoint( joint, B, true, a.p.h, a.p.inv_h );
			// joint events: colored joints only, as the reference's overflow pass does not test
			if ( ( joint->forceThreshold < FLT_MAX || joint->torqueThreshold < FLT_MAX ) && joint->eventFlag == 0 &&
				 JointReactionReached( joint, a.p.inv_h ) )
			{
				joint->eventFlag = 1;
			}
		}
		else if ( Stage == kStageRelax )
		{
			SolveJoint( joint, B, false, a.p.h, a.p.inv_h );
		}
		return;
	}

	int k = t - r.jointCount;
	if ( k >= r.contactCount )
	{
		return;
	}
	int i = r.contactStart + k;

	if ( Stage == kStageWarmStart )
	{
		WarmStartContact( a.c, i, B );
	}
	else if ( Stage == kStageSolve )
	{
		SolveContact( a.c, i, B, true, a.p.inv_h, -a.p.contactSpeed, false );
	}
	else if ( Stage == kStageRelax )
	{
		bool rolling = GroupAnyNonZero( a.c, kcRollingResistance, r.contactStart, r.contactCount, k );
		SolveContact( a.c, i, B, false, a.p.inv_h, -a.p.contactSpeed, rolling );
	}
	else
	{
		if ( GroupAnyNonZero( a.c, kcRestitution, r.contactStart, r.contactCount, k ) )
		{
			RestitutionContact( a.c, i, B, a.p.restitutionThreshold );
		}
	}
}

// Overflow constraints of one stage, walked by a single lane: joints then contacts in order, as
// the reference's single-threaded overflow passes do (src/solver.c:1100-1168).
template <int Stage, typename Bodies>
B2G_DEV void OverflowSerial( const ColorArgs& a, const ColorRange& r, const Bodies& B )
{
	for ( int t = 0; t < r.jointCount; ++t )
	{
		JointSim* joint = a.joints + a.jointList[r.jointStart + t];
		if ( Stage == kStageWarmStart )
		{
			WarmStartJoint( joint, B );
		}
		else if ( Stage == kStageSolve )
		{
			SolveJoint( joint, B, true, a.p.h, a.p.inv_h );
		}
		else if ( Stage == kStageRelax )
		{
			SolveJoint( joint, B, false, a.p.h, a.p.inv_h );
		}
	}

	for ( int k = 0; k < r.contactCount; ++k )
	{
		int i = r.contactStart + k;
		if ( Stage == kStageWarmStart )
		{
			OverflowWarmStart( a.c, i, B );
		}
		else if ( Stage == kStageSolve )
		{
			OverflowSolve( a.c, i, B, true, a.p.inv_h, a.p.contactSpeed );
		}
		else if ( Stage == kStageRelax )
		{
			OverflowSolve( a.c, i, B, false, a.p.inv_h, a.p.contactSpeed );
		}
		else
		{
			OverflowRestitution( a.c, i, B, a.p.restitutionThreshold );
		}
	}
}

// One graph color of one stage over the id-indexed body states in device memory.
template <int Stage>
__global__ void ColorKernel( ColorArgs a )
{
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	ColorRange r = { a.jointStart, a.jointCount, a.contactStart, a.contactCount };
	ColorLane<Stage>( a, r, t, GlobalBodies{ a.states } );
}

// Consecutive small colors of one stage in one workgroup, separated by workgroup barriers instead
// of kernel boundaries (each color's lanes and arithmetic are those of ColorKernel).
constexpr int kTailThreads = 1024;
constexpr int kTailMaxColors = 24;

struct TailArgs
{
	ColorArgs color;
	int colorCount;
	int jointStart[kTailMaxColors + 1];
	int contactStart[kTailMaxColors + 1];
};

template <int Stage>
__global__ __launch_bounds__( kTailThreads ) void ColorTailKernel( TailArgs a )
{
	const GlobalBodies B{ a.color.states };
	for ( int c = 0; c < a.colorCount; ++c )
	{
		ColorRange r;
		r.jointStart = a.jointStart[c];
		r.jointCount = a.jointStart[c + 1] - r.jointStart;
		r.contactStart = a.contactStart[c];
		r.contactCount = a.contactStart[c + 1] - r.contactStart;
		int lanes = r.jointCount + r.contactCount;
		if ( lanes == 0 || ( Stage == kStageRestitution && r.contactCount == 0 ) )
		{
			continue;
		}
		for ( int t = threadIdx.x; t < lanes; t += blockDim.x )
		{
			ColorLane<Stage>( a.color, r, t, B );
		}
		__syncthreads();
	}
}

template <int Stage>
__global__ void OverflowKernel( ColorArgs a )
{
	if ( blockIdx.x != 0 || threadIdx.x != 0 )
	{
		return;
	}
	ColorRange r = { a.jointStart, a.jointCount, a.contactStart, a.contactCount };
	OverflowSerial<Stage>( a, r, GlobalBodies{ a.states } );
}

// Overflow constraints by level (SolverLists::overflowPlan), one workgroup. The overflow joints go
// first, walked in order by one lane as in OverflowSerial. Then each chunk's constraint
// rows and body states are copied into LDS, the levels run one after another with a workgroup
// barrier between them (a level's contacts share no dynamic body), and the rows and dynamic
// velocities go back to device memory. The per-contact arithmetic is OverflowSerial's; only
// contacts that touch disjoint state run at the same time.
constexpr int kOverflowThreads = 256;

constexpr size_t OverflowLdsBytes()
{
	return (size_t)kOverflowChunkContacts * ( kcFieldCount * sizeof( float ) + 6 * sizeof( int ) ) + sizeof( int ) +
		   (size_t)kOverflowChunkBodies * ( 7 * sizeof( float ) + sizeof( uint16_t ) );
}

static_assert( OverflowLdsBytes() <= 160 * 1024, "overflow chunk must fit in LDS" );

template <int Stage>
__global__ __launch_bounds__( kOverflowThreads ) void OverflowLevelKernel( ColorArgs a, const int* plan )
{
	extern __shared__ float lds[];
	const int C = kOverflowChunkContacts;
	const int NB = kOverflowChunkBodies;
	float* fields = lds;
	int* ints = reinterpret_cast<int*>( fields + kcFieldCount * C );
	int* sEntries = ints + 4 * C;	  // the chunk's entries (chunk-local contact index) by level
	int* sLevels = sEntries + C;	  // the chunk's level bounds, relative to its first entry
	float* bf = reinterpret_cast<float*>( sLevels + C + 1 );
	uint16_t* bflags = reinterpret_cast<uint16_t*>( bf + 7 * NB );

	ConstraintSoA lc = a.c;
	lc.f = FieldTable{ fields, C };
	lc.bodyA = ints;
	lc.bodyB = ints + C;
	lc.pointCount = ints + 2 * C;
	lc.contactId = ints + 3 * C;
	const LdsBodies B{ bf, bf + NB, bf + 2 * NB, bf + 3 * NB, bf + 4 * NB, bf + 5 * NB, bf + 6 * NB, bflags };
	const GlobalBodies G{ a.states };

	if ( a.jointCount > 0 && Stage != kStageRestitution )
	{
		if ( threadIdx.x == 0 )
		{
			ColorRange r = { a.jointStart, a.jointCount, a.contactStart, 0 };
			OverflowSerial<Stage>( a, r, G );
		}
		__syncthreads();
	}

	const int chunkCount = plan[0];
	const int* bodies = plan + plan[1];
	const int* slots = plan + plan[2];
	const int* entries = plan + plan[3];
	const int* levels = plan + plan[4];
	const int tid = threadIdx.x;

	for ( int ch = 0; ch < chunkCount; ++ch )
	{
		const int* t = plan + 5 + 6 * ch;
		int c0 = t[0], c1 = t[1], b0 = t[2], b1 = t[3], l0 = t[4], l1 = t[5];
		int n = c1 - c0;
		int first = a.contactStart + c0;

		for ( int x = tid; x < kcFieldCount * n; x += kOverflowThreads )
		{
			int k = x / n;
			int j = x - k * n;
			fields[k * C + j] = a.c.f[k][first + j];
		}
		for ( int j = tid; j < n; j += kOverflowThreads )
		{
			ints[j] = slots[2 * ( c0 + j )];
			ints[C + j] = slots[2 * ( c0 + j ) + 1];
			ints[2 * C + j] = a.c.pointCount[first + j];
			ints[3 * C + j] = a.c.contactId[first + j];
			sEntries[j] = entries[c0 + j];
		}
		for ( int l = tid; l <= l1 - l0; l += kOverflowThreads )
		{
			sLevels[l] = levels[l0 + l] - c0;
		}
		for ( int j = tid; j < b1 - b0; j += kOverflowThreads )
		{
			LaneBody lb = G.Load( bodies[b0 + j] );
			B.vx[j] = lb.vx;
			B.vy[j] = lb.vy;
			B.w[j] = lb.w;
			B.dpx[j] = lb.dpx;
			B.dpy[j] = lb.dpy;
			B.dqc[j] = lb.dqc;
			B.dqs[j] = lb.dqs;
			B.flags[j] = (uint16_t)lb.flags;
		}
		__syncthreads();

		auto solve = [&]( int j ) {
			if ( Stage == kStageWarmStart )
			{
				OverflowWarmStart( lc, j, B );
			}
			else if ( Stage == kStageSolve )
			{
				OverflowSolve( lc, j, B, true, a.p.inv_h, a.p.contactSpeed );
			}
			else if ( Stage == kStageRelax )
			{
				OverflowSolve( lc, j, B, false, a.p.inv_h, a.p.contactSpeed );
			}
			else
			{
				OverflowRestitution( lc, j, B, a.p.restitutionThreshold );
			}
		};
		const int levelCount = l1 - l0;
		int l = 0;
		while ( l < levelCount )
		{
			if ( sLevels[l + 1] - sLevels[l] == 1 )
			{
				// a run of one-contact levels (a chain through one body) goes to one lane, with no
				// barrier between its levels
				int r = l + 1;
				while ( r < levelCount && sLevels[r + 1] - sLevels[r] == 1 )
				{
					++r;
				}
				if ( tid == 0 )
				{
					for ( int m = l; m < r; ++m )
					{
						solve( sEntries[sLevels[m]] );
					}
				}
				l = r;
			}
			else
			{
				for ( int e = sLevels[l] + tid; e < sLevels[l + 1]; e += kOverflowThreads )
				{
					solve( sEntries[e] );
				}
				l += 1;
			}
			__syncthreads();
		}

		for ( int x = tid; x < kcFieldCount * n; x += kOverflowThreads )
		{
			int k = x / n;
			int j = x - k * n;
			a.c.f[k][first + j] = fields[k * C + j];
		}
		for ( int j = tid; j < b1 - b0; j += kOverflowThreads )
		{
			LaneBody lb;
			lb.vx = B.vx[j];
			lb.vy = B.vy[j];
			lb.w = B.w[j];
			lb.flags = B.flags[j];
			G.Store( bodies[b0 + j], lb );
		}
		__syncthreads();
	}
}

// ---------------------------------------------------------------------------------------------
// Fused solver: the whole substep loop of one world in one workgroup.
//
// The world's awake body states live in LDS for the duration of the solve (7 floats and 16 flag
// bits per body, 30 bytes, so kFusedBodyCapacity bodies fit in 160 KB). Colors and stages are
// separated by workgroup barriers instead of kernel boundaries; constraint fields stream from
// device memory. Stage order and per-constraint arithmetic are those of the color-kernel path
// (reference src/solver.c:1090-1170).

constexpr int kFusedThreads = 1024;
constexpr int kFusedMaxContacts = 4096;
constexpr int kFusedMaxOverflow = 8;

// Automatic solver path choice: the fused solver for worlds with at most kFusedMaxContacts
// contact constraints and at most kFusedMaxOverflow overflow constraints.
static bool AutoFused( int constraintCount, int overflowCount )
{
	return constraintCount <= kFusedMaxContacts && overflowCount <= kFusedMaxOverflow;
}
constexpr int kMaxColorSlots = 24;

struct FusedArgs
{
	ColorArgs color; // c (with slot arrays), joints, jointList, p; ranges set per color below
	const int* awake;
	int bodyCount;
	const BodySim* sims;
	BodyState* states;
	int activeColorCount;
	int colorContactStart[kMaxColorSlots + 1];
	int colorJointStart[kMaxColorSlots + 1];
	int overflowContactStart;
	int overflowContactCount;
	int overflowJointStart;
	int overflowJointCount;
	int subStepCount;
	int anyRestitution;
	const int* warmStart; // warm start by body (WarmStartBody), indexed by awake index
	const int* warmList;
	int warmByBody; // 1: warm start body by body, 0: color by color
};

template <int Stage>
B2G_DEV void FusedStage( const FusedArgs& a, const LdsBodies& B )
{
	if ( a.overflowContactCount > 0 || a.overflowJointCount > 0 )
	{
		if ( threadIdx.x == 0 )
		{
			ColorRange r = { a.overflowJointStart, a.overflowJointCount, a.overflowContactStart, a.overflowContactCount };
			OverflowSerial<Stage>( a.color, r, B );
		}
		__syncthreads();
	}

	for ( int color = 0; color < a.activeColorCount; ++color )
	{
		ColorRange r;
		r.jointStart = a.colorJointStart[color];
		r.jointCount = a.colorJointStart[color + 1] - r.jointStart;
		r.contactStart = a.colorContactStart[color];
		r.contactCount = a.colorContactStart[color + 1] - r.contactStart;
		int lanes = r.jointCount + r.contactCount;
		if ( lanes == 0 || ( Stage == kStageRestitution && r.contactCount == 0 ) )
		{
			continue;
		}
		for ( int t = threadIdx.x; t < lanes; t += blockDim.x )
		{
			ColorLane<Stage>( a.color, r, t, B );
		}
		__syncthreads();
	}
}

// LDS bytes of a world with `count` awake bodies: slot 0 holds the static identity state, slots
// 1..count the awake bodies; 7 floats and 16 flag bits per slot.
__host__ __device__ inline int FusedLdsBytes( int count )
{
	int cap = count + 1;
	return ( 7 * 4 * cap + 2 * cap + 15 ) & ~15;
}

// The whole substep loop of one world by the calling workgroup (any block size), body states in
// the dynamic LDS at lds.
B2G_DEV void FusedBody( const FusedArgs& a, char* lds )
{
	const int n = a.bodyCount;
	const int cap = n + 1;
	float* s_vx = reinterpret_cast<float*>( lds );
	float* s_vy = s_vx + cap;
	float* s_w = s_vy + cap;
	float* s_dpx = s_w + cap;
	float* s_dpy = s_dpx + cap;
	float* s_dqc = s_dpy + cap;
	float* s_dqs = s_dqc + cap;
	uint16_t* s_flags = reinterpret_cast<uint16_t*>( s_dqs + cap );

	LdsBodies B{ s_vx, s_vy, s_w, s_dpx, s_dpy, s_dqc, s_dqs, s_flags };
	const GlobalBodies G{ a.states };

	auto put = [&]( int slot, const BodyState& st ) {
		s_vx[slot] = st.linearVelocity.x;
		s_vy[slot] = st.linearVelocity.y;
		s_w[slot] = st.angularVelocity;
		s_flags[slot] = (uint16_t)st.flags;
		s_dpx[slot] = st.deltaPosition.x;
		s_dpy[slot] = st.deltaPosition.y;
		s_dqc[slot] = st.deltaRotation.c;
		s_dqs[slot] = st.deltaRotation.s;
	};

	for ( int i = threadIdx.x; i < n; i += blockDim.x )
	{
		put( i + 1, ToState( G.Load( a.awake[i] ) ) );
	}
	if ( threadIdx.x == 0 )
	{
		// identity state of static bodies (reference b2_identityBodyState)
		BodyState id;
		id.linearVelocity = { 0.0f, 0.0f };
		id.angularVelocity = 0.0f;
		id.flags = 0;
		id.deltaPosition = { 0.0f, 0.0f };
		id.deltaRotation = { 1.0f, 0.0f };
		put( kStaticSlot, id );
	}
	__syncthreads();

	for ( int sub = 0; sub < a.subStepCount; ++sub )
	{
		// velocity integration, then the warm-start stage color by color or body by body (one
		// barrier instead of one per color; the floats are the same either way)
		for ( int i = threadIdx.x; i < n; i += blockDim.x )
		{
			BodyState st = ToState( B.Load( i + 1 ) );
			IntegrateVelocity( a.sims[a.awake[i]], &st, a.color.p );
			float vx = st.linearVelocity.x, vy = st.linearVelocity.y, w = st.angularVelocity;
			if ( a.warmByBody )
			{
				WarmStartBody( a.color.c, a.color.joints, a.color.jointList, a.overflowContactStart, a.warmList, a.warmStart[i],
							   a.warmStart[i + 1], B, st.deltaRotation, st.deltaPosition, vx, vy, w );
			}
			s_vx[i + 1] = vx;
			s_vy[i + 1] = vy;
			s_w[i + 1] = w;
		}
		__syncthreads();

		if ( a.warmByBody == 0 )
		{
			FusedStage<kStageWarmStart>( a, B );
		}
		FusedStage<kStageSolve>( a, B );

		for ( int i = threadIdx.x; i < n; i += blockDim.x )
		{
			BodyState st = ToState( B.Load( i + 1 ) );
			IntegratePosition( &st, a.color.p );
			put( i + 1, st );
		}
		__syncthreads();

		FusedStage<kStageRelax>( a, B );
	}

	if ( a.anyRestitution )
	{
		FusedStage<kStageRestitution>( a, B );
	}

	// write the whole state back (flags carry the speed-cap bit to finalize)
	for ( int i = threadIdx.x; i < n; i += blockDim.x )
	{
		BodyState* dst = a.states + a.awake[i];
		dst->linearVelocity = { s_vx[i + 1], s_vy[i + 1] };
		dst->angularVelocity = s_w[i + 1];
		dst->flags = ( dst->flags & 0xFFFF0000u ) | s_flags[i + 1];
		dst->deltaPosition = { s_dpx[i + 1], s_dpy[i + 1] };
		dst->deltaRotation = { s_dqc[i + 1], s_dqs[i + 1] };
	}
}

__global__ __launch_bounds__( kFusedThreads ) void SolveFusedKernel( FusedArgs a )
{
	extern __shared__ char s_lds[];
	FusedBody( a, s_lds );
}

// Persistent form over many worlds: each workgroup takes the next world from a queue (worlds
// ordered largest first) until the queue is empty. Every workgroup reaches the queue's end, so
// the launch drains without any inter-group wait.
__global__ __launch_bounds__( kFusedThreads ) void SolveFusedBatchKernel( const FusedArgs* args, const int* order, int count,
																		  int* queue )
{
	extern __shared__ char s_lds[];
	__shared__ int s_world;
	for ( ;; )
	{
		if ( threadIdx.x == 0 )
		{
			s_world = atomicAdd( queue, 1 );
		}
		__syncthreads();
		// the world index is uniform across the workgroup: keep it (and the argument loads) scalar
		int w = __builtin_amdgcn_readfirstlane( s_world );
		if ( w >= count )
		{
			break;
		}
		int world = __builtin_amdgcn_readfirstlane( order[w] );
		FusedBody( args[world], s_lds );
		__syncthreads();
	}
}

// ---------------------------------------------------------------------------------------------
// XCD solver: the whole substep loop of one world in one launch, on the compute units of one XCD.
//
// A stage's colors are Gauss-Seidel phases, and in the color-kernel path each phase is a kernel
// launch with a floor of 3-5 us. Here the phases are separated by a barrier among the workgroups
// of one XCD instead. The CUs of an XCD share one L2, so a barrier only has to make each
// workgroup's stores reach that L2 before it arrives (a wait for its outstanding memory
// operations) and count arrivals with L2 atomics: about 1.2-1.4 us per barrier on MI355X, where
// coherence across XCDs would mean writing back each L2 at every barrier. The launch places
// workgroup b on XCD b mod 8 (round robin), so the grid is 8 x kXcdGroups workgroups
// (cooperative: all resident at once); the workgroups that read XCC_ID 0 take tickets and solve,
// the others exit at once. Every wait is bounded: a barrier that is not complete after
// kXcdSpinLimit polls (or sees another workgroup's timeout) raises the error flag and the
// workgroup exits, and the host stops with an error.
//
// What workgroups share must not be read from a CU's L1, which may still hold a line another CU
// has since rewritten (measured on MI355X: plain loads read stale data, and an L1 invalidate
// costs about 0.9 us per workgroup per barrier because it also walks the L2). Body states, the
// only mutable data the phases exchange, are therefore read with agent-scope loads
// (CoherentBodies), which skip the L1. Constraint and joint rows are otherwise only touched by the
// wave that owns them: a phase's lanes are cut into 64-lane waves and wave w goes to participant
// w mod kXcdGroups. The per-body warm start (worlds without joints) is the exception: it reads
// other waves' impulses with the same kind of loads and leaves the total-normal-impulse sums to
// each contact's own solve lane (AddWarmTotals); worlds with joints warm start color by color.
// Each phase runs exactly the lanes one color kernel launch runs, in any order within the phase.

constexpr int kXcdThreads = 1024;
constexpr int kXcdGroups = 32; // workgroups on the solving XCD (one per CU of an MI355X XCD)
constexpr int kXcdCount = 8;
constexpr int kXcdGrid = kXcdGroups * kXcdCount;
constexpr int kXcdWaves = kXcdThreads / 64;
constexpr long long kXcdSpinLimit = 1 << 22; // polls before a barrier gives up (~seconds)

// sync[]: 0 ticket, 1 barrier arrivals, 2 error (barrier timeout), 3 unused
struct XcdArgs
{
	FusedArgs f; // color ranges, warm lists and parameters; f.color.c holds body ids
	int* sync;
	int zero; // 0, passed at run time so the barrier's polling atomic is not folded into a load
};

__device__ __forceinline__ int XccId()
{
	// XCC_ID hardware register (4 bits)
	return __builtin_amdgcn_s_getreg( 20 | ( 0 << 6 ) | ( 3 << 11 ) ) & 0xF;
}

__device__ __forceinline__ int L2Poll( int* p, int zero )
{
	return __hip_atomic_fetch_add( p, zero, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP );
}

// Barrier among the participants; returns false when it timed out (or another one did).
__device__ bool XcdBarrier( const XcdArgs& a, int& arrivals )
{
	__shared__ int s_ok;
	arrivals += kXcdGroups;
	__builtin_amdgcn_s_waitcnt( 0 ); // this wave's stores are in L2
	__syncthreads();
	if ( threadIdx.x == 0 )
	{
		__hip_atomic_fetch_add( a.sync + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP );
		int ok = 1;
		long long polls = 0;
		while ( L2Poll( a.sync + 1, a.zero ) < arrivals )
		{
			if ( ++polls > kXcdSpinLimit || L2Poll( a.sync + 2, a.zero ) != 0 )
			{
				__hip_atomic_fetch_add( a.sync + 2, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP );
				ok = 0;
				break;
			}
		}
		s_ok = ok;
	}
	__syncthreads();
	return s_ok != 0;
}

// Lanes of one color of one stage, wave-interleaved over the participants.
template <int Stage>
B2G_DEV void XcdColor( const XcdArgs& a, const ColorRange& r, int rank, const CoherentBodies& B )
{
	int lanes = r.jointCount + r.contactCount;
	int lane = threadIdx.x & 63;
	for ( int lw = threadIdx.x >> 6;; lw += kXcdWaves )
	{
		int w = lw * kXcdGroups + rank;
		if ( w * 64 >= lanes )
		{
			break;
		}
		int t = w * 64 + lane;
		if ( Stage == kStageSolve && a.f.warmByBody && t >= r.jointCount && t < lanes )
		{
			AddWarmTotals( a.f.color.c, r.contactStart + t - r.jointCount );
		}
		ColorLane<Stage>( a.f.color, r, t, B );
	}
}

template <int Stage>
B2G_DEV bool XcdStage( const XcdArgs& a, int rank, const CoherentBodies& B, int& arrivals )
{
	for ( int color = 0; color < a.f.activeColorCount; ++color )
	{
		ColorRange r;
		r.jointStart = a.f.colorJointStart[color];
		r.jointCount = a.f.colorJointStart[color + 1] - r.jointStart;
		r.contactStart = a.f.colorContactStart[color];
		r.contactCount = a.f.colorContactStart[color + 1] - r.contactStart;
		int lanes = r.jointCount + r.contactCount;
		if ( lanes == 0 || ( Stage == kStageRestitution && r.contactCount == 0 ) )
		{
			continue;
		}
		XcdColor<Stage>( a, r, rank, B );
		if ( XcdBarrier( a, arrivals ) == false )
		{
			return false;
		}
	}
	return true;
}

__global__ __launch_bounds__( kXcdThreads ) void SolveXcdKernel( XcdArgs a )
{
	__shared__ int s_rank;
	if ( threadIdx.x == 0 )
	{
		int rank = -1;
		if ( XccId() == 0 )
		{
			rank = __hip_atomic_fetch_add( a.sync, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP );
		}
		s_rank = rank < kXcdGroups ? rank : -1;
	}
	__syncthreads();
	const int rank = s_rank;
	if ( rank < 0 )
	{
		return;
	}

	const FusedArgs& f = a.f;
	const CoherentBodies B{ f.states };
	const StepParams& p = f.color.p;
	const int n = f.bodyCount;
	const int lane = threadIdx.x & 63;
	int arrivals = 0;

	for ( int sub = 0; sub < f.subStepCount; ++sub )
	{
		for ( int lw = threadIdx.x >> 6;; lw += kXcdWaves )
		{
			int w = lw * kXcdGroups + rank;
			if ( w * 64 >= n )
			{
				break;
			}
			int i = w * 64 + lane;
			if ( i < n )
			{
				int id = f.awake[i];
				BodyState st = B.LoadState( id );
				IntegrateVelocity( f.sims[id], &st, p );
				float vx = st.linearVelocity.x, vy = st.linearVelocity.y, wv = st.angularVelocity;
				if ( f.warmByBody )
				{
					WarmStartBody<4, false, true>( f.color.c, f.color.joints, f.color.jointList, f.overflowContactStart, f.warmList,
												   f.warmStart[i], f.warmStart[i + 1], B, st.deltaRotation, st.deltaPosition, vx, vy, wv );
				}
				f.states[id].linearVelocity = { vx, vy };
				f.states[id].angularVelocity = wv;
			}
		}
		if ( XcdBarrier( a, arrivals ) == false )
		{
			return;
		}
		if ( f.warmByBody == 0 && XcdStage<kStageWarmStart>( a, rank, B, arrivals ) == false )
		{
			return;
		}
		if ( XcdStage<kStageSolve>( a, rank, B, arrivals ) == false )
		{
			return;
		}
		for ( int lw = threadIdx.x >> 6;; lw += kXcdWaves )
		{
			int w = lw * kXcdGroups + rank;
			if ( w * 64 >= n )
			{
				break;
			}
			int i = w * 64 + lane;
			if ( i < n )
			{
				int id = f.awake[i];
				BodyState st = B.LoadState( id );
				IntegratePosition( &st, p );
				f.states[id] = st;
			}
		}
		if ( XcdBarrier( a, arrivals ) == false )
		{
			return;
		}
		if ( XcdStage<kStageRelax>( a, rank, B, arrivals ) == false )
		{
			return;
		}
	}
	if ( f.anyRestitution )
	{
		XcdStage<kStageRestitution>( a, rank, B, arrivals );
	}
}

struct StoreArgs
{
	ConstraintSoA c;
	ContactSim* contacts;
	int coloredCount;
	int totalCount;
	int* hitCandidates; // contact ids of colored contacts flagged for a hit event
	int* hitCount;
	float negHitThreshold; // -hitEventThreshold
};

// Impulses back to the manifolds (reference src/contact_solver.c:2238-2331 for colored
// constraints, which write both points, and :516-545 for overflow, which write pointCount points).
B2G_DEV void StoreImpulsesBody( const StoreArgs& a, int i )
{
	if ( i >= a.totalCount )
	{
		return;
	}

	const ConstraintSoA& c = a.c;
	b2Manifold* m = &a.contacts[c.contactId[i]].manifold;
	int writeCount = i < a.coloredCount ? 2 : m->pointCount;

	for ( int j = 0; j < writeCount; ++j )
	{
		int o = j == 0 ? 0 : ( kcAnchorA2X - kcAnchorA1X );
		m->points[j].normalImpulse = c.f[kcNormalImpulse1 + o][i];
		m->points[j].tangentImpulse = c.f[kcTangentImpulse1 + o][i];
		m->points[j].totalNormalImpulse = c.f[kcTotalNormalImpulse1 + o][i];
		m->points[j].normalVelocity = c.f[kcRelativeVelocity1 + o][i];
	}
	m->rollingImpulse = c.f[kcRollingImpulse][i];

	// hit-event candidates, colored contacts only (reference src/contact_solver.c:2305-2319); the
	// point total impulse rules out speculative points that did not collide
	if ( i < a.coloredCount && ( a.contacts[c.contactId[i]].simFlags & kSimEnableHitEvent ) != 0 )
	{
		for ( int k = 0; k < m->pointCount; ++k )
		{
			const b2ManifoldPoint* mp = m->points + k;
			if ( mp->normalVelocity < a.negHitThreshold && mp->totalNormalImpulse > 0.0f )
			{
				a.hitCandidates[atomicAdd( a.hitCount, 1 )] = c.contactId[i];
				break;
			}
		}
	}
}

// Fat shape AABB: float world points, double extremes and inflation, one narrowing cast
// (reference src/geometry.c:402-505).
B2G_DEV b2AABB FatShapeAABB( const ShapeDev& s, b2Transform xf, float extra )
{
	switch ( s.type )
	{
		case b2_capsuleShape:
		{
			b2Vec2 v1 = b2TransformWorldPoint( xf, s.capsule.center1 );
			b2Vec2 v2 = b2TransformWorldPoint( xf, s.capsule.center2 );
			double r = (double)s.capsule.radius + (double)extra;
			double v1x = v1.x, v1y = v1.y, v2x = v2.x, v2y = v2.y;
			b2AABB aabb = { { (float)( ( v1x < v2x ? v1x : v2x ) - r ), (float)( ( v1y < v2y ? v1y : v2y ) - r ) },
							{ (float)( ( v1x > v2x ? v1x : v2x ) + r ), (float)( ( v1y > v2y ? v1y : v2y ) + r ) } };
			return aabb;
		}
		case b2_circleShape:
		{
			b2Vec2 c = b2TransformWorldPoint( xf, s.circle.center );
			double r = (double)s.circle.radius + (double)extra;
			double cx = c.x, cy = c.y;
			b2AABB aabb = { { (float)( cx - r ), (float)( cy - r ) }, { (float)( cx + r ), (float)( cy + r ) } };
			return aabb;
		}
		case b2_polygonShape:
		{
			b2Vec2 v = b2TransformWorldPoint( xf, s.polygon.vertices[0] );
			double lx = v.x, ly = v.y, ux = v.x, uy = v.y;
			for ( int i = 1; i < s.polygon.count; ++i )
			{
				v = b2TransformWorldPoint( xf, s.polygon.vertices[i] );
				double vx = v.x, vy = v.y;
				lx = vx < lx ? vx : lx;
				ly = vy < ly ? vy : ly;
				ux = vx > ux ? vx : ux;
				uy = vy > uy ? vy : uy;
			}
			double r = (double)s.polygon.radius + (double)extra;
			b2AABB aabb = { { (float)( lx - r ), (float)( ly - r ) }, { (float)( ux + r ), (float)( uy + r ) } };
			return aabb;
		}
		case b2_segmentShape:
		case b2_chainSegmentShape:
		{
			const b2Segment& seg = s.type == b2_segmentShape ? s.segment : s.chainSegment.segment;
			b2Vec2 v1 = b2TransformWorldPoint( xf, seg.point1 );
			b2Vec2 v2 = b2TransformWorldPoint( xf, seg.point2 );
			double e = (double)extra;
			double v1x = v1.x, v1y = v1.y, v2x = v2.x, v2y = v2.y;
			b2AABB aabb = { { (float)( ( v1x < v2x ? v1x : v2x ) - e ), (float)( ( v1y < v2y ? v1y : v2y ) - e ) },
							{ (float)( ( v1x > v2x ? v1x : v2x ) + e ), (float)( ( v1y > v2y ? v1y : v2y ) + e ) } };
			return aabb;
		}
		default:
		{
			b2AABB empty = { { 0.0f, 0.0f }, { 0.0f, 0.0f } };
			return empty;
		}
	}
}

struct FinalizeArgs
{
	const int* awake;
	int count;
	BodySim* sims;
	BodyState* states;
	BodyAux* aux;
	ShapeDev* shapes;
	StepParams p;
	int* counters; // [0] enlarged bodies, [1] enlarged shapes, [2] fast bodies
	int* fastBodies;
	int* enlargedBodies;
	int* enlargedShapes;
	b2AABB* enlargedShapeAABBs;
	// island sleep (null when sleep is off): per awake index the island's slot in the awake set,
	// negative when the island has removed constraints, and the island id
	const int* islandSlots;
	const int* islandIds;
	uint32_t* islandBits;		  // awake islands, by slot
	unsigned long long* splitKey; // sleepiest split candidate: sleep time bits << 32 | island id
	// the pinned mailbox's enlarged-body, enlarged-shape and box lists, written beside the device
	// lists (null: the pack copies the device lists instead)
	int* mailBodies;
	int* mailShapes;
	float* mailAabbs;
};

// Body finalization (reference src/solver.c:553-743).
B2G_DEV void FinalizeBodiesBody( const FinalizeArgs& a, int simIndex )
{
	if ( simIndex >= a.count )
	{
		return;
	}

	int bodyId = a.awake[simIndex];
	BodyState* state = a.states + bodyId;
	BodySim* sim = a.sims + bodyId;
	BodyAux* aux = a.aux + bodyId;

	b2Vec2 v = state->linearVelocity;
	float w = state->angularVelocity;

	sim->center = b2OffsetPos( sim->center, state->deltaPosition );
	sim->transform.q = b2NormalizeRot( b2MulRot( state->deltaRotation, sim->transform.q ) );

	float maxVelocity = b2Length( v ) + b2AbsFloat( w ) * sim->maxExtent;
	float maxDeltaPosition = b2Length( state->deltaPosition ) + b2AbsFloat( state->deltaRotation.s ) * sim->maxExtent;
	float positionSleepFactor = 0.5f;
	float sleepVelocity = b2MaxFloat( maxVelocity, positionSleepFactor * a.p.inv_dt * maxDeltaPosition );

	state->deltaPosition = { 0.0f, 0.0f };
	state->deltaRotation = { 1.0f, 0.0f };

	sim->transform.p = b2OffsetPos( sim->center, b2Neg( b2RotateVector( sim->transform.q, sim->localCenter ) ) );

	sim->force = { 0.0f, 0.0f };
	sim->torque = 0.0f;

	uint32_t bodyFlags = aux->flags;
	bodyFlags &= ~kTransientFlags;
	bodyFlags |= ( sim->flags & ( kIsSpeedCapped | kHadTimeOfImpact ) );
	bodyFlags |= ( state->flags & ( kIsSpeedCapped | kHadTimeOfImpact ) );
	sim->flags &= ~kTransientFlags;
	state->flags &= ~kTransientFlags;

	float sleepTime = aux->sleepTime;
	if ( a.p.enableSleep == 0 || ( bodyFlags & kEnableSleep ) == 0 || sleepVelocity > aux->sleepThreshold )
	{
		sleepTime = 0.0f;
		float maxMotion = b2MaxFloat( maxDeltaPosition, maxVelocity * a.p.dt );
		if ( aux->type == b2_dynamicBody && a.p.enableContinuous && maxMotion > aux->safetyFactor * sim->minExtent )
		{
			bodyFlags |= kIsFast;
			int slot = atomicAdd( a.counters + 2, 1 );
			a.fastBodies[slot] = simIndex;
		}
		else
		{
			sim->center0 = sim->center;
			sim->rotation0 = sim->transform.q;
		}
	}
	else
	{
		sim->center0 = sim->center;
		sim->rotation0 = sim->transform.q;
		sleepTime += a.p.dt;
	}

	aux->sleepTime = sleepTime;
	aux->flags = bodyFlags;
	// Any body of an island keeps it awake; a sleepy body of an island with removed constraints is
	// a split candidate, the sleepiest winning and ties going to the larger island id
	// (reference src/solver.c:676-694).
	if ( a.islandBits != nullptr )
	{
		int slot = a.islandSlots[simIndex];
		if ( sleepTime < B2_TIME_TO_SLEEP )
		{
			int index = slot & 0x7fffffff;
			atomicOr( a.islandBits + ( index >> 5 ), 1u << ( index & 31 ) );
		}
		else if ( slot < 0 )
		{
			unsigned long long key = (unsigned long long)__float_as_uint( sleepTime ) << 32 | (uint32_t)a.islandIds[simIndex];
			atomicMax( a.splitKey, key );
		}
	}

	b2Transform transform = sim->transform;
	bool isFast = ( bodyFlags & kIsFast ) != 0;
	bool enlarged = false;
	int shapeId = sim->headShapeId;
	while ( shapeId != B2_NULL_INDEX )
	{
		ShapeDev* shape = a.shapes + shapeId;
		if ( isFast )
		{
			enlarged = true;
		}
		else
		{
			b2AABB aabb = FatShapeAABB( *shape, transform, kSpeculativeDistance );
			shape->aabb = aabb;
			shape->enlargedAABB = 0;
			if ( b2AABB_Contains( shape->fatAABB, aabb ) == false )
			{
				float margin = shape->aabbMargin;
				b2AABB fatAABB;
				fatAABB.lowerBound.x = aabb.lowerBound.x - margin;
				fatAABB.lowerBound.y = aabb.lowerBound.y - margin;
				fatAABB.upperBound.x = aabb.upperBound.x + margin;
				fatAABB.upperBound.y = aabb.upperBound.y + margin;
				shape->fatAABB = fatAABB;
				shape->enlargedAABB = 1;
				enlarged = true;
				int slot = atomicAdd( a.counters + 1, 1 );
				a.enlargedShapes[slot] = shapeId;
				a.enlargedShapeAABBs[slot] = fatAABB;
				if ( a.mailShapes != nullptr )
				{
					a.mailShapes[slot] = shapeId;
					float* dst = a.mailAabbs + 4 * slot;
					dst[0] = fatAABB.lowerBound.x;
					dst[1] = fatAABB.lowerBound.y;
					dst[2] = fatAABB.upperBound.x;
					dst[3] = fatAABB.upperBound.y;
				}
			}
		}
		shapeId = shape->nextShapeId;
	}

	aux->enlarged = enlarged ? 1 : 0;
	if ( enlarged )
	{
		int slot = atomicAdd( a.counters + 0, 1 );
		a.enlargedBodies[slot] = simIndex;
		if ( a.mailBodies != nullptr )
		{
			a.mailBodies[slot] = simIndex;
		}
	}
}

// ---------------------------------------------------------------------------------------------
// Continuous collision of fast bodies against static shapes (reference src/solver.c:181-550).
//
// One lane per fast non-bullet body. The lane walks a device copy of the host's static tree in
// the reference's query order (stack, child2 popped first; src/dynamic_tree.c:1085-1135). The
// order matters: each time of impact is bounded by the smallest fraction found so far, and the
// root finder's iterates depend on that bound. All geometry runs in a frame centered on the
// body's start position (the reference's base frame).

// Static tree node: fat AABB, category bits (union over the subtree), children or leaf shape id.
struct CcdNode
{
	b2AABB aabb;
	uint64_t categoryBits;
	int child1;
	int child2;
	int shapeId; // leaf only
	int leaf;
};

constexpr int kCcdStackSize = 256;

struct ContinuousArgs
{
	const int* fastBodies;
	const int* fastCount;
	const int* awake;
	BodySim* sims;
	ShapeDev* shapes;
	const CcdNode* tree;
	int treeRoot;
	int* error; // set when a query stack would overflow
	const int2* noCollide; // (body, static body) pairs joined with collideConnected off
	int noCollideCount;
	const uint64_t* filterRejects; // sorted (static shape << 32 | fast shape) keys the custom filter rejected
	int filterRejectCount;
	int* enlargedShapeCount;
	int* enlargedShapes;
	b2AABB* enlargedShapeAABBs;
	float huge;
	IdPair* sensorHits; // (sensor shape, fast shape) hits on static sensors; null when the world has none
	int* sensorHitCount;
	int sensorHitCapacity;
	int* mailShapes; // as FinalizeArgs::mailShapes
	float* mailAabbs;
};

constexpr int kMaxContinuousSensorHits = 8;
constexpr int kSensorHitCountSlot = 9;	   // device counter slots
constexpr int kSplitKeySlot = 14;		   // 64-bit, slots 14 and 15
constexpr int kSensorOverlapCountSlot = 10;
constexpr int kFinishDoneSlot = 11; // workgroups of FinishKernel done (the last one resets it)
constexpr int kCollideDoneSlot = 3;	 // workgroups of CollideAndPackKernel done (likewise)

B2G_DEV bool ShouldShapesCollide( const b2Filter& a, const b2Filter& b )
{
	if ( a.groupIndex == b.groupIndex && a.groupIndex != 0 )
	{
		return a.groupIndex > 0;
	}
	return ( a.maskBits & b.categoryBits ) != 0 && ( a.categoryBits & b.maskBits ) != 0;
}

// Distance proxy of a shape in its body frame (reference src/shape.c:1055-1076).
B2G_DEV b2ShapeProxy ShapeProxy( const ShapeDev& s )
{
	switch ( s.type )
	{
		case b2_capsuleShape:
			return MakeProxy( &s.capsule.center1, 2, s.capsule.radius );
		case b2_circleShape:
			return MakeProxy( &s.circle.center, 1, s.circle.radius );
		case b2_polygonShape:
			return MakeProxy( s.polygon.vertices, s.polygon.count, s.polygon.radius );
		case b2_segmentShape:
			return MakeProxy( &s.segment.point1, 2, 0.0f );
		default:
			return MakeProxy( &s.chainSegment.segment.point1, 2, 0.0f );
	}
}

// Local centroid (reference src/shape.c:705-722) and smallest extent about it (src/shape.c:826-880).
B2G_DEV b2Vec2 ShapeCentroidDev( const ShapeDev& s )
{
	switch ( s.type )
	{
		case b2_capsuleShape:
			return b2Lerp( s.capsule.center1, s.capsule.center2, 0.5f );
		case b2_circleShape:
			return s.circle.center;
		case b2_polygonShape:
			return s.polygon.centroid;
		case b2_segmentShape:
			return b2Lerp( s.segment.point1, s.segment.point2, 0.5f );
		default:
			return b2Lerp( s.chainSegment.segment.point1, s.chainSegment.segment.point2, 0.5f );
	}
}

B2G_DEV float ShapeMinExtent( const ShapeDev& s, float huge )
{
	switch ( s.type )
	{
		case b2_capsuleShape:
			return s.capsule.radius;
		case b2_circleShape:
			return s.circle.radius;
		case b2_polygonShape:
		{
			float minExtent = huge;
			for ( int i = 0; i < s.polygon.count; ++i )
			{
				b2Vec2 v = s.polygon.vertices[i];
				float planeOffset = b2Dot( s.polygon.normals[i], b2Sub( v, s.polygon.centroid ) );
				minExtent = b2MinFloat( minExtent, planeOffset );
			}
			return minExtent + s.polygon.radius;
		}
		default:
			return 0.0f;
	}
}

// AABB translated from the base frame to world: double sums narrowed once (reference src/aabb.h:55-63).
B2G_DEV b2AABB OffsetAABB( b2AABB box, b2Vec2 origin )
{
	b2AABB r;
	r.lowerBound.x = (float)( origin.x + (double)box.lowerBound.x );
	r.lowerBound.y = (float)( origin.y + (double)box.lowerBound.y );
	r.upperBound.x = (float)( origin.x + (double)box.upperBound.x );
	r.upperBound.y = (float)( origin.y + (double)box.upperBound.y );
	return r;
}

B2G_DEV b2Sweep RelativeSweep( const BodySim& sim, b2Vec2 base )
{
	b2Sweep sw;
	sw.c1 = b2SubPos( sim.center0, base );
	sw.c2 = b2SubPos( sim.center, base );
	sw.q1 = sim.rotation0;
	sw.q2 = sim.transform.q;
	sw.localCenter = sim.localCenter;
	return sw;
}

constexpr float kCoreFraction = 0.25f;

B2G_DEV void EnlargeShape( const ContinuousArgs& a, int shapeId, ShapeDev* shape )
{
	if ( b2AABB_Contains( shape->fatAABB, shape->aabb ) == false )
	{
		float margin = shape->aabbMargin;
		b2AABB fat;
		fat.lowerBound.x = shape->aabb.lowerBound.x - margin;
		fat.lowerBound.y = shape->aabb.lowerBound.y - margin;
		fat.upperBound.x = shape->aabb.upperBound.x + margin;
		fat.upperBound.y = shape->aabb.upperBound.y + margin;
		shape->fatAABB = fat;
		shape->enlargedAABB = 1;
		int slot = atomicAdd( a.enlargedShapeCount, 1 );
		a.enlargedShapes[slot] = shapeId;
		a.enlargedShapeAABBs[slot] = fat;
		if ( a.mailShapes != nullptr )
		{
			a.mailShapes[slot] = shapeId;
			float* dst = a.mailAabbs + 4 * slot;
			dst[0] = fat.lowerBound.x;
			dst[1] = fat.lowerBound.y;
			dst[2] = fat.upperBound.x;
			dst[3] = fat.upperBound.y;
		}
	}
}

B2G_DEV bool FilterRejected( const ContinuousArgs& a, int staticShapeId, int fastShapeId )
{
	uint64_t key = (uint64_t)(uint32_t)staticShapeId << 32 | (uint32_t)fastShapeId;
	int lo = 0;
	int hi = a.filterRejectCount;
	while ( lo < hi )
	{
		int mid = ( lo + hi ) >> 1;
		if ( a.filterRejects[mid] < key )
		{
			lo = mid + 1;
		}
		else
		{
			hi = mid;
		}
	}
	return lo < a.filterRejectCount && a.filterRejects[lo] == key;
}

B2G_DEV void ContinuousBodyAt( const ContinuousArgs& a, int bodyId );

B2G_DEV void ContinuousBody( const ContinuousArgs& a, int k )
{
	if ( k >= *a.fastCount )
	{
		return;
	}
	ContinuousBodyAt( a, a.awake[a.fastBodies[k]] );
}

// The sweep of one fast body (by body id), after its finalization.
B2G_DEV void ContinuousBodyAt( const ContinuousArgs& a, int bodyId )
{
	BodySim* fastSim = a.sims + bodyId;
	if ( fastSim->flags & kIsBullet )
	{
		// bullets sweep after the refit, against the moving bodies too (the host does them)
		return;
	}

	b2Vec2 base = fastSim->center0;
	b2Sweep sweep = RelativeSweep( *fastSim, base );

	b2Transform xf1;
	xf1.q = sweep.q1;
	xf1.p = b2Sub( sweep.c1, b2RotateVector( sweep.q1, sweep.localCenter ) );
	b2Transform xf2;
	xf2.q = sweep.q2;
	xf2.p = b2Sub( sweep.c2, b2RotateVector( sweep.q2, sweep.localCenter ) );

	float fraction = 1.0f;
	bool hadImpact = false;
	// sensor hits no later than the solid hit found so far (reference src/solver.c:319-335)
	IdPair sensorHits[kMaxContinuousSensorHits];
	float sensorFractions[kMaxContinuousSensorHits];
	int sensorCount = 0;

	for ( int shapeId = fastSim->headShapeId; shapeId != B2_NULL_INDEX; shapeId = a.shapes[shapeId].nextShapeId )
	{
		ShapeDev* fastShape = a.shapes + shapeId;
		fastShape->enlargedAABB = 0;
		b2Vec2 centroid1 = b2TransformPoint( xf1, fastShape->localCentroid );
		b2Vec2 centroid2 = b2TransformPoint( xf2, fastShape->localCentroid );

		b2AABB box1 = fastShape->aabb;
		b2AABB box2 = OffsetAABB( FatShapeAABB( *fastShape, xf2, 0.0f ), base );
		fastShape->aabb = box2;
		// no continuous collision for sensors, but their bounds follow (reference src/solver.c:436-440)
		if ( fastShape->flags & kShapeIsSensor )
		{
			continue;
		}
		b2AABB swept = b2AABB_Union( box1, box2 );

		int stack[kCcdStackSize];
		int stackCount = 0;
		if ( a.treeRoot != B2_NULL_INDEX )
		{
			stack[stackCount++] = a.treeRoot;
		}

		while ( stackCount > 0 )
		{
			const CcdNode& node = a.tree[stack[--stackCount]];
			if ( b2AABB_Overlaps( node.aabb, swept ) == false || node.categoryBits == 0 )
			{
				continue;
			}
			if ( node.leaf == 0 )
			{
				if ( stackCount < kCcdStackSize - 1 )
				{
					stack[stackCount++] = node.child1;
					stack[stackCount++] = node.child2;
				}
				else
				{
					*a.error = 1;
				}
				continue;
			}

			int otherId = node.shapeId;
			const ShapeDev& other = a.shapes[otherId];
			if ( otherId == shapeId || other.bodyId == fastShape->bodyId )
			{
				continue;
			}
			// a sensor counts only when both shapes enable sensor events (reference src/solver.c:209-214)
			const bool otherSensor = ( other.flags & kShapeIsSensor ) != 0;
			if ( otherSensor && ( ( other.flags & fastShape->flags & kShapeEnableSensorEvents ) == 0 || a.sensorHits == nullptr ) )
			{
				continue;
			}
			if ( ShouldShapesCollide( fastShape->filter, other.filter ) == false )
			{
				continue;
			}
			bool jointed = false;
			for ( int n = 0; n < a.noCollideCount; ++n )
			{
				if ( a.noCollide[n].x == bodyId && a.noCollide[n].y == other.bodyId )
				{
					jointed = true;
				}
			}
			if ( jointed )
			{
				continue;
			}
			// custom filtering (reference src/solver.c:242-256), asked on the host before the step
			if ( a.filterRejectCount > 0 && FilterRejected( a, otherId, shapeId ) )
			{
				continue;
			}

			const BodySim& otherSim = a.sims[other.bodyId];

			if ( other.type == b2_chainSegmentShape )
			{
				// skip fast parallel motion over a one-sided chain (reference src/solver.c:265-290)
				b2Transform rel = { b2SubPos( otherSim.transform.p, base ), otherSim.transform.q };
				b2Vec2 p1 = b2TransformPoint( rel, other.chainSegment.segment.point1 );
				b2Vec2 p2 = b2TransformPoint( rel, other.chainSegment.segment.point2 );
				float length;
				b2Vec2 e = b2GetLengthAndNormalize( &length, b2Sub( p2, p1 ) );
				if ( length > kLinearSlop )
				{
					float separation1 = b2Cross( b2Sub( centroid1, p1 ), e );
					float separation2 = b2Cross( b2Sub( centroid2, p1 ), e );
					float coreDistance = kCoreFraction * fastSim->minExtent;
					if ( separation1 < 0.0f || ( separation1 - separation2 < coreDistance && separation2 > coreDistance ) )
					{
						continue;
					}
				}
			}

			b2TOIInput input;
			input.proxyA = ShapeProxy( other );
			input.proxyB = ShapeProxy( *fastShape );
			input.sweepA = RelativeSweep( otherSim, base );
			input.sweepB = sweep;
			input.maxFraction = fraction;

			b2TOIOutput output = TimeOfImpact( input );
			if ( otherSensor )
			{
				if ( output.fraction <= fraction && sensorCount < kMaxContinuousSensorHits )
				{
					sensorHits[sensorCount] = IdPair{ otherId, shapeId };
					sensorFractions[sensorCount] = output.fraction;
					sensorCount += 1;
				}
				continue;
			}
			float hitFraction = fraction;
			bool didHit = false;
			if ( 0.0f < output.fraction && output.fraction < fraction )
			{
				hitFraction = output.fraction;
				didHit = true;
			}
			else if ( 0.0f == output.fraction )
			{
				// fall back to a small circle at the fast shape's centroid
				b2Vec2 centroid = ShapeCentroidDev( *fastShape );
				float radius = kCoreFraction * ShapeMinExtent( *fastShape, a.huge );
				input.proxyB = MakeProxy( &centroid, 1, radius );
				output = TimeOfImpact( input );
				if ( 0.0f < output.fraction && output.fraction < fraction )
				{
					hitFraction = output.fraction;
					didHit = true;
				}
			}

			if ( didHit )
			{
				hadImpact = true;
				fraction = hitFraction;
			}
		}
	}

	if ( hadImpact )
	{
		fastSim->flags |= kHadTimeOfImpact;
	}

	// sensor hits that happened before the final solid hit (reference src/solver.c:540-547)
	for ( int i = 0; i < sensorCount; ++i )
	{
		if ( sensorFractions[i] < fraction )
		{
			int slot = atomicAdd( a.sensorHitCount, 1 );
			if ( slot < a.sensorHitCapacity )
			{
				a.sensorHits[slot] = sensorHits[i];
			}
		}
	}

	if ( fraction < 1.0f )
	{
		b2Rot q = b2NLerp( sweep.q1, sweep.q2, fraction );
		b2Vec2 c = b2Lerp( sweep.c1, sweep.c2, fraction );
		b2Vec2 origin = b2Sub( c, b2RotateVector( q, sweep.localCenter ) );

		fastSim->transform.q = q;
		fastSim->transform.p = b2OffsetPos( base, origin );
		fastSim->center = b2OffsetPos( base, c );
		fastSim->rotation0 = q;
		fastSim->center0 = fastSim->center;

		b2Transform transform = fastSim->transform;
		for ( int shapeId = fastSim->headShapeId; shapeId != B2_NULL_INDEX; shapeId = a.shapes[shapeId].nextShapeId )
		{
			ShapeDev* shape = a.shapes + shapeId;
			shape->aabb = FatShapeAABB( *shape, transform, kSpeculativeDistance );
			EnlargeShape( a, shapeId, shape );
		}
	}
	else
	{
		fastSim->rotation0 = fastSim->transform.q;
		fastSim->center0 = fastSim->center;
		for ( int shapeId = fastSim->headShapeId; shapeId != B2_NULL_INDEX; shapeId = a.shapes[shapeId].nextShapeId )
		{
			EnlargeShape( a, shapeId, a.shapes + shapeId );
		}
	}
}

// ---------------------------------------------------------------------------------------------
// Lane kernels: each per-item stage body runs either as a single-world launch or as one launch
// over many worlds, where a block table maps every block to (world, block within the world).

// ---------------------------------------------------------------------------------------------
// Step records for the adjoint: after constraint preparation and before the substep loop, one
// lane per awake body and per contact constraint copies what the backward pass recomputes the
// step from (see b2amd_TapeStep, b2amd_TapeBody, b2amd_TapeConstraint).

size_t TapeConstraintsOffset( int bodyCount )
{
	return kTapeBodiesOffset + ( ( (size_t)bodyCount * sizeof( b2amd_TapeBody ) + 255 ) & ~(size_t)255 );
}

size_t TapeRecordBytes( int bodyCount, int constraintCount )
{
	return ( ( TapeConstraintsOffset( bodyCount ) + (size_t)constraintCount * sizeof( b2amd_TapeConstraint ) + 255 ) &
			 ~(size_t)255 );
}

struct TapeRecordArgs
{
	b2amd_TapeStep header;
	char* dst;
	size_t constraintsOffset;
	ConstraintSoA c; // bodyA/bodyB are body ids, slotA/slotB fused solver slots
	const ContactSim* contacts;
	const BodySim* sims;
	const BodyState* states;
	const int* awake;
	int bodyCount;
	int constraintCount;
};

B2G_DEV void TapeRecordBody( const TapeRecordArgs& a, int t )
{
	if ( t == 0 )
	{
		*reinterpret_cast<b2amd_TapeStep*>( a.dst ) = a.header;
	}

	if ( t < a.bodyCount )
	{
		int id = a.awake[t];
		const BodySim& sim = a.sims[id];
		const BodyState& st = a.states[id];
		b2amd_TapeBody b;
		b.bodyId = id;
		b.flags = st.flags;
		b.center = sim.center;
		b.origin = sim.transform.p;
		b.rotation = sim.transform.q;
		b.linearVelocity = st.linearVelocity;
		b.angularVelocity = st.angularVelocity;
		b.invMass = sim.invMass;
		b.invInertia = sim.invInertia;
		b.gravityScale = sim.gravityScale;
		b.linearDamping = sim.linearDamping;
		b.angularDamping = sim.angularDamping;
		b.localCenter = sim.localCenter;
		b.force = sim.force;
		b.torque = sim.torque;
		b.padding = 0.0f;
		reinterpret_cast<b2amd_TapeBody*>( a.dst + kTapeBodiesOffset )[t] = b;
	}

	if ( t < a.constraintCount )
	{
		const ConstraintSoA& c = a.c;
		int contactId = c.contactId[t];
		const ContactSim& cs = a.contacts[contactId];
		const b2Manifold& m = cs.manifold;
		b2amd_TapeConstraint r;
		r.contactId = contactId;
		r.shapeIdA = cs.shapeIdA;
		r.shapeIdB = cs.shapeIdB;
		r.bodyIdA = c.bodyA[t];
		r.bodyIdB = c.bodyB[t];
		r.slotA = c.slotA[t];
		r.slotB = c.slotB[t];
		r.pointCount = m.pointCount;
		r.info = cs.stepInfo;
		r.normal = m.normal;
		for ( int j = 0; j < 2; ++j )
		{
			int o = j == 0 ? 0 : ( kcAnchorA2X - kcAnchorA1X );
			bool live = j < m.pointCount;
			r.id[j] = live ? m.points[j].id : 0;
			r.anchorA[j] = live ? m.points[j].anchorA : b2Vec2{ 0.0f, 0.0f };
			r.anchorB[j] = live ? m.points[j].anchorB : b2Vec2{ 0.0f, 0.0f };
			r.separation[j] = live ? m.points[j].separation : 0.0f;
			r.baseSeparation[j] = live ? m.points[j].baseSeparation : 0.0f;
			r.normalImpulse[j] = c.f[kcNormalImpulse1 + o][t];
			r.tangentImpulse[j] = c.f[kcTangentImpulse1 + o][t];
		}
		r.friction = c.f[kcFriction][t];
		r.restitution = c.f[kcRestitution][t];
		r.rollingResistance = c.f[kcRollingResistance][t];
		r.tangentSpeed = c.f[kcTangentSpeed][t];
		r.biasRate = c.f[kcBiasRate][t];
		r.massScale = c.f[kcMassScale][t];
		r.impulseScale = c.f[kcImpulseScale][t];
		r.cachedRotationA = cs.cachedRotationA;
		r.cachedRotationB = cs.cachedRotationB;
		r.padding = 0.0f;
		reinterpret_cast<b2amd_TapeConstraint*>( a.dst + a.constraintsOffset )[t] = r;
	}
}

struct TapeRecordOp
{
	using Args = TapeRecordArgs;
	static B2G_DEV void Run( const Args& a, int i )
	{
		TapeRecordBody( a, i );
	}
};

struct CollideOp
{
	using Args = CollideArgs;
	static B2G_DEV void Run( const Args& a, int i )
	{
		CollideBody( a, i, DeviceCollideSink{ a } );
	}
};

struct PrepareContactsOp
{
	using Args = PrepareArgs;
	static B2G_DEV void Run( const Args& a, int i )
	{
		PrepareContactsBody( a, i );
	}
};

struct PrepareJointsOp
{
	using Args = PrepareJointArgs;
	static B2G_DEV void Run( const Args& a, int i )
	{
		PrepareJointsBody( a, i );
	}
};

struct StoreImpulsesOp
{
	using Args = StoreArgs;
	static B2G_DEV void Run( const Args& a, int i )
	{
		StoreImpulsesBody( a, i );
	}
};

struct FinalizeBodiesOp
{
	using Args = FinalizeArgs;
	static B2G_DEV void Run( const Args& a, int i )
	{
		FinalizeBodiesBody( a, i );
	}
};

struct ContinuousOp
{
	using Args = ContinuousArgs;
	static B2G_DEV void Run( const Args& a, int i )
	{
		ContinuousBody( a, i );
	}
};

struct InitContactsOp
{
	using Args = InitArgs;
	static B2G_DEV void Run( const Args& a, int i )
	{
		InitContactsBody( a, i );
	}
};

// Staged host-to-device copies of a batch: each workgroup copies one entry, word by word, from
// the batch's uploaded staging buffer to the entry's destination.
struct ScatterEntry
{
	uint32_t* dst;
	uint64_t srcOffset; // bytes into the staging buffer
	uint32_t words;
	uint32_t pad;
};

__global__ void ScatterKernel( const ScatterEntry* entries, const char* base )
{
	ScatterEntry e = entries[blockIdx.x];
	const uint32_t* src = reinterpret_cast<const uint32_t*>( base + e.srcOffset );
	for ( uint32_t w = threadIdx.x; w < e.words; w += blockDim.x )
	{
		e.dst[w] = src[w];
	}
}

template <typename Op>
__global__ void LaneKernel( typename Op::Args a )
{
	Op::Run( a, blockIdx.x * blockDim.x + threadIdx.x );
}

template <typename Op>
__global__ void LaneBatchKernel( const typename Op::Args* args, const int2* blocks )
{
	int2 b = blocks[blockIdx.x];
	Op::Run( args[b.x], b.y * blockDim.x + threadIdx.x );
}

// ---------------------------------------------------------------------------------------------
// Kernel classes for timing, with the algorithmic bytes each unit of work must move: every field
// a lane reads or writes counted once, body states gathered and scattered per constraint
// (per-unit figures of the step's data layout).

enum KernelClass
{
	kkCollide,
	kkPrepareContacts,
	kkPrepareJoints,
	kkIntegrateVelocities,
	kkWarmStart,
	kkSolve,
	kkIntegratePositions,
	kkRelax,
	kkRestitution,
	kkOverflow,
	kkStore,
	kkFinalize,
	kkSolveFused,
	kkContinuous,
	kkColorTail,
	kkSolveXcd,
	kkFinish,
	kKernelClassCount
};

static const char* s_kernelClassNames[kKernelClassCount] = {
	"collide", "prepare_contacts", "prepare_joints", "integrate_velocities_warm_start", "warm_start", "solve",
	"integrate_positions", "relax", "restitution", "overflow", "store_impulses", "finalize_bodies", "solve_fused", "continuous",
	"color_tail", "solve_xcd", "finish",
};

constexpr double kBytesCollideFull = 800.0;		// per contact with a new manifold
constexpr double kBytesCollideRecycled = 180.0; // per recycled contact
constexpr double kBytesPrepare = 292.0;			// per contact constraint
constexpr double kBytesIntegrateVelocities = 56.0;
constexpr double kBytesIntegratePositions = 60.0;
constexpr double kBytesWarmStart = 148.0;
constexpr double kBytesSolve = 212.0;
constexpr double kBytesRelax = 236.0;
constexpr double kBytesRestitution = 236.0;
constexpr double kBytesStore = 76.0;
constexpr double kBytesFinalize = 370.0;
constexpr double kBytesJoint = 256.0; // revolute joint record read and written plus two body states
// Fused solver (body states in LDS): constraint fields of warm start 100 B, solve 124 B and relax
// 156 B per contact per substep; 32 B of body sim per body per substep; 64 B of body state read
// and written once per step.
constexpr double kBytesFusedContactSubstep = 380.0;
constexpr double kBytesFusedBodySubstep = 32.0;
constexpr double kBytesFusedBodyStep = 64.0;

// ---------------------------------------------------------------------------------------------
// Readback: one small kernel per host sync point copies counters and variable-length result
// lists into a pinned host mailbox and clears the counters for the next step, so each sync
// point costs one launch and one stream synchronize.

constexpr int kMailHeader = 8;
// Device counter slot that keeps the last solve's enlarged-shape count (the pack kernel clears
// the finalize counters it reads).
constexpr int kKeptShapeCountSlot = 12;
constexpr int kPackThreads = 1024;

struct PackCollideArgs
{
	int* counters;
	const int* changes;
	int* mail;
	int deferredOffset; // changes and mailbox offset of the deferred contact ids (2 per collided contact)
	int listsInMail;	// the narrowphase lanes wrote the lists into the mailbox already
};

B2G_DEV void PackCollideBody( const PackCollideArgs& a )
{
	int count = a.counters[0];
	int recycled = a.counters[1];
	int deferred = a.counters[2];
	for ( int i = threadIdx.x; a.listsInMail == 0 && i < count; i += blockDim.x )
	{
		a.mail[kMailHeader + i] = a.changes[i];
	}
	for ( int i = threadIdx.x; a.listsInMail == 0 && i < deferred; i += blockDim.x )
	{
		a.mail[kMailHeader + a.deferredOffset + i] = a.changes[a.deferredOffset + i];
	}
	__syncthreads();
	if ( threadIdx.x == 0 )
	{
		a.mail[0] = count;
		a.mail[1] = recycled;
		a.mail[2] = deferred;
		a.counters[0] = 0;
		a.counters[1] = 0;
		a.counters[2] = 0;
	}
}

__global__ __launch_bounds__( kPackThreads ) void PackCollideKernel( PackCollideArgs a )
{
	PackCollideBody( a );
}

// The narrowphase of a single world with its pack: the last workgroup to finish (each workgroup
// releases its writes at device scope before it counts itself done; the last one acquires)
// copies the change records into the mailbox, writes the counts and clears the counters. No
// workgroup waits for another.
__global__ void CollideAndPackKernel( CollideArgs a, PackCollideArgs pc, int* done )
{
	CollideOp::Run( a, blockIdx.x * blockDim.x + threadIdx.x );
	// The barrier completes every thread's stores (they reach the L2); thread 0's device-scope
	// fence then releases the whole workgroup's writes before it counts the workgroup done.
	__shared__ int s_last;
	__syncthreads();
	if ( threadIdx.x == 0 )
	{
		__threadfence();
		s_last = atomicAdd( done, 1 ) == (int)gridDim.x - 1 ? 1 : 0;
	}
	__syncthreads();
	if ( s_last == 0 )
	{
		return;
	}
	__threadfence();
	PackCollideBody( pc );
	if ( threadIdx.x == 0 )
	{
		*done = 0;
	}
}

// one workgroup per world
__global__ __launch_bounds__( kPackThreads ) void PackCollideBatchKernel( const PackCollideArgs* args )
{
	PackCollideBody( args[blockIdx.x] );
}

// One hit event in the readback mailbox: contact id, shape ids, world point, normal and approach
// speed (floats as bits).
constexpr int kHitRecordInts = 8;

struct PackFinalizeArgs
{
	// hit events, built after finalize and continuous collision moved the bodies
	const ContactSim* contacts;
	const BodySim* sims;
	const BodyAux* aux;
	const int* hitCandidates;
	int* hitCount;
	float hitThreshold;
	int hitOffset; // mailbox offset in ints
	int* counters; // enlarged bodies, enlarged shapes, fast bodies, continuous-collision error
	const int* enlargedBodies;
	const int* enlargedShapes;
	const b2AABB* enlargedShapeAABBs;
	uint32_t* islandBits;		  // awake islands (null when sleep is off): copied, then cleared
	unsigned long long* splitKey; // likewise
	int islandWords;
	int* keptShapeCount;	 // receives the enlarged-shape count before it is cleared (may be null)
	const int* xcdSync;		 // XCD solver barrier words of this step (null when it did not run)
	JointSim* joints;		 // joint events: flagged joints are listed and their flags cleared
	int jointCount;			 // 0 when no joint has a finite threshold
	int jointEventOffset;	 // mailbox offset in ints
	int bodyCount;
	int shapeOffset; // mailbox offsets in ints
	int aabbOffset;
	int sleepOffset;
	int* mail;
	int listsInMail; // finalize and the sweeps wrote the enlarged lists into the mailbox already
};

B2G_DEV void PackFinalizeBody( const PackFinalizeArgs& a )
{
	int bodies = a.counters[0];
	int shapes = a.counters[1];
	int fast = a.counters[2];
	int error = a.counters[3];
	for ( int i = threadIdx.x; a.listsInMail == 0 && i < bodies; i += blockDim.x )
	{
		a.mail[kMailHeader + i] = a.enlargedBodies[i];
	}
	for ( int i = threadIdx.x; a.listsInMail == 0 && i < shapes; i += blockDim.x )
	{
		a.mail[a.shapeOffset + i] = a.enlargedShapes[i];
		b2AABB box = a.enlargedShapeAABBs[i];
		float* dst = reinterpret_cast<float*>( a.mail + a.aabbOffset + 4 * i );
		dst[0] = box.lowerBound.x;
		dst[1] = box.lowerBound.y;
		dst[2] = box.upperBound.x;
		dst[3] = box.upperBound.y;
	}
	if ( a.islandBits != nullptr )
	{
		uint32_t* dst = reinterpret_cast<uint32_t*>( a.mail + a.sleepOffset );
		for ( int i = threadIdx.x; i < a.islandWords; i += blockDim.x )
		{
			dst[i] = a.islandBits[i];
			a.islandBits[i] = 0;
		}
		if ( threadIdx.x == 0 )
		{
			unsigned long long key = *a.splitKey;
			dst[a.islandWords] = (uint32_t)key;
			dst[a.islandWords + 1] = (uint32_t)( key >> 32 );
			*a.splitKey = 0;
		}
	}

	// Joint events (reference src/solver.c:1648-1700), listed in any order (the host sorts them)
	__shared__ int jointEvents;
	if ( threadIdx.x == 0 )
	{
		jointEvents = 0;
	}
	__syncthreads();
	for ( int i = threadIdx.x; i < a.jointCount; i += blockDim.x )
	{
		if ( a.joints[i].eventFlag != 0 )
		{
			a.joints[i].eventFlag = 0;
			int slot = atomicAdd( &jointEvents, 1 );
			a.mail[a.jointEventOffset + slot] = i;
		}
	}

	// Hit events (reference src/solver.c:1703-1820): the fastest approaching point that carried
	// impulse; the world point follows a static body's anchor when the other body is not static.
	int hits = *a.hitCount;
	for ( int h = threadIdx.x; h < hits; h += blockDim.x )
	{
		int contactId = a.hitCandidates[h];
		const ContactSim& cs = a.contacts[contactId];
		float approach = a.hitThreshold;
		int best = -1;
		for ( int p = 0; p < cs.manifold.pointCount; ++p )
		{
			const b2ManifoldPoint* mp = cs.manifold.points + p;
			float approachSpeed = -mp->normalVelocity;
			if ( approachSpeed > approach && mp->totalNormalImpulse > 0.0f )
			{
				approach = approachSpeed;
				best = p;
			}
		}
		int* rec = a.mail + a.hitOffset + kHitRecordInts * h;
		rec[0] = best >= 0 ? contactId : -1;
		rec[1] = cs.shapeIdA;
		rec[2] = cs.shapeIdB;
		if ( best >= 0 )
		{
			const b2ManifoldPoint* mp = cs.manifold.points + best;
			b2Pos point;
			if ( a.aux[cs.bodyIdA].type != b2_staticBody && a.aux[cs.bodyIdB].type == b2_staticBody )
			{
				point = b2OffsetPos( a.sims[cs.bodyIdB].center, mp->anchorB );
			}
			else
			{
				point = b2OffsetPos( a.sims[cs.bodyIdA].center, mp->anchorA );
			}
			rec[3] = __float_as_int( point.x );
			rec[4] = __float_as_int( point.y );
			rec[5] = __float_as_int( cs.manifold.normal.x );
			rec[6] = __float_as_int( cs.manifold.normal.y );
			rec[7] = __float_as_int( approach );
		}
	}
	__syncthreads();
	if ( threadIdx.x == 0 )
	{
		a.mail[4] = hits;
		*a.hitCount = 0;
		a.mail[0] = bodies;
		a.mail[1] = shapes;
		a.mail[2] = fast;
		a.mail[3] = error;
		if ( a.keptShapeCount != nullptr )
		{
			*a.keptShapeCount = shapes;
		}
		a.mail[5] = a.xcdSync != nullptr ? a.xcdSync[2] : 0;
		a.mail[6] = jointEvents;
		a.counters[0] = 0;
		a.counters[1] = 0;
		a.counters[2] = 0;
		a.counters[3] = 0;
	}
}

__global__ __launch_bounds__( kPackThreads ) void PackFinalizeKernel( PackFinalizeArgs a )
{
	PackFinalizeBody( a );
}

__global__ __launch_bounds__( kPackThreads ) void PackFinalizeBatchKernel( const PackFinalizeArgs* args )
{
	PackFinalizeBody( args[blockIdx.x] );
}

// The end of a single world's solve in one launch: the first storeBlocks workgroups store the
// contact impulses, the others finalize the awake bodies, each fast non-bullet body sweeping
// right after its own finalization (its sweep reads only static shapes and its own state). The
// last workgroup to finish packs the results for the host. Each workgroup releases its writes at
// device scope before it counts itself done, and the last workgroup acquires before it reads, so
// the pack sees every other workgroup's results; no workgroup waits for another.
struct FinishArgs
{
	StoreArgs store;
	FinalizeArgs fin;
	ContinuousArgs ccd;
	PackFinalizeArgs pack;
	int storeBlocks;
	int sweep; // run the continuous sweeps
	int* done;
};

// 64-lane workgroups: the bodies spread over as many compute units as possible (1024-thread
// workgroups measured at twice the time of 256); the last one's pack has little left to copy
constexpr int kFinishThreads = 64;

__global__ __launch_bounds__( kFinishThreads ) void FinishKernel( FinishArgs a )
{
	if ( (int)blockIdx.x < a.storeBlocks )
	{
		int i = blockIdx.x * blockDim.x + threadIdx.x;
		if ( i < a.store.totalCount )
		{
			StoreImpulsesBody( a.store, i );
		}
	}
	else
	{
		int i = ( blockIdx.x - a.storeBlocks ) * blockDim.x + threadIdx.x;
		FinalizeBodiesBody( a.fin, i );
		if ( a.sweep && i < a.fin.count )
		{
			int bodyId = a.fin.awake[i];
			if ( ( a.fin.aux[bodyId].flags & kIsFast ) != 0 && ( a.fin.sims[bodyId].flags & kIsBullet ) == 0 )
			{
				ContinuousBodyAt( a.ccd, bodyId );
			}
		}
	}

	// The barrier completes every thread's stores (they reach the L2); thread 0's device-scope
	// fence then releases the whole workgroup's writes before it counts the workgroup done.
	__shared__ int s_last;
	__syncthreads();
	if ( threadIdx.x == 0 )
	{
		__threadfence();
		s_last = atomicAdd( a.done, 1 ) == (int)gridDim.x - 1 ? 1 : 0;
	}
	__syncthreads();
	if ( s_last == 0 )
	{
		return;
	}
	__threadfence();
	PackFinalizeBody( a.pack );
	if ( threadIdx.x == 0 )
	{
		*a.done = 0;
	}
}

// ---------------------------------------------------------------------------------------------
// Device buffers

template <typename T>
struct DevArray
{
	T* ptr = nullptr;
	size_t capacity = 0;

	void Reserve( size_t n, bool keep, hipStream_t stream, size_t used )
	{
		if ( n <= capacity )
		{
			return;
		}
		// doubling with a floor of 256 elements keeps reallocations (allocation, copy, free: each
		// a synchronizing runtime call) rare while a world's contact set grows
		size_t newCap = std::max<size_t>( n, 2 * capacity );
		newCap = std::max<size_t>( newCap, 256 );
		T* p = nullptr;
		HIP_CHECK( hipMalloc( &p, newCap * sizeof( T ) ) );
		if ( keep && ptr != nullptr && used > 0 )
		{
			HIP_CHECK( hipMemcpyAsync( p, ptr, used * sizeof( T ), hipMemcpyDeviceToDevice, stream ) );
			HIP_CHECK( hipStreamSynchronize( stream ) );
		}
		if ( ptr != nullptr )
		{
			HIP_CHECK( hipFree( ptr ) );
		}
		ptr = p;
		capacity = newCap;
	}

	void Free()
	{
		if ( ptr != nullptr )
		{
			(void)hipFree( ptr );
		}
		ptr = nullptr;
		capacity = 0;
	}
};

struct DeviceWorld::Impl
{
	hipStream_t stream = nullptr;

	DevArray<BodySim> sims;
	DevArray<BodyState> states;
	DevArray<BodyAux> aux;
	int bodyCount = 0;

	DevArray<ShapeDev> shapes;
	int shapeCount = 0;

	DevArray<ContactSim> contacts;
	int contactCapacityUsed = 0;

	DevArray<JointSim> joints;
	int jointCount = 0;

	DevArray<ContactInit> inits;

	// work lists
	DevArray<int> awakeBodies;
	DevArray<int> bodySlot; // body id -> fused solver LDS slot (awake index), kStaticSlot otherwise
	DevArray<int> collideList;
	DevArray<int> jointList;
	DevArray<int> warmStart; // per awake body: its range of warmList (warm start by body)
	DevArray<int> warmList;
	DevArray<int> overflowPlan;
	DevArray<int> changes;
	SolverLists lists;
	std::vector<int> slotStaging;
	int solverPath = 0;
	// step recording for the adjoint (null when not recording)
	TapeStore* tape = nullptr;
	int tapeIndex = 0;
	// Workgroup size of the color kernels (64..1024; launches are latency-bound, so it matters little).
	int colorBlock = ColorBlockFromEnv();

	// constraints
	DevArray<float> cfields;
	DevArray<int> cints;
	size_t constraintCapacity = 0;
	ConstraintSoA soa = {};

	// finalize outputs
	DevArray<int> enlargedBodies;
	DevArray<int> fastBodies;
	DevArray<CcdNode> ccdTree;
	int ccdTreeRoot = B2_NULL_INDEX;
	DevArray<int2> ccdNoCollide;
	DevArray<IdPair> sensorHits;	  // continuous collision hits on static sensors
	DevArray<IdPair> sensorOverlaps; // the sensor stage's overlaps
	DevArray<int> sensorShapes;
	int sensorHitCapacity = 0;
	int sensorOverlapCapacity = 0;
	DevArray<uint64_t> ccdFilterRejects;
	int ccdFilterRejectCount = 0;
	std::vector<CcdNode> ccdTreeStaging;
	std::vector<int2> ccdNoCollideStaging;
	DevArray<int> enlargedShapes;
	DevArray<b2AABB> enlargedShapeAABBs;
	DevArray<int> islandSlots;
	DevArray<int> islandIds;
	DevArray<uint32_t> islandBits; // kept zero between steps by the pack kernel
	size_t islandBitsZeroed = 0;   // words known zero
	int islandCount = 0;
	DevArray<int> hitCandidates;
	int solveHitCapacity = 0;

	// counters: [0..1] collide, [4..6] finalize
	int* counters = nullptr;
	int* hostCounters = nullptr;

	// Batched stepping: instead of launching, the phase functions record each kernel's
	// arguments here and a batch launches one kernel per stage over all worlds.
	bool recording = false;
	struct Record
	{
		bool collide = false, prepare = false, prepareJoints = false, fused = false, store = false, finalize = false,
			 continuous = false, packCollide = false, packFinalize = false, tapeRecord = false;
		TapeRecordArgs tapeArgs;
		int tapeLanes = 0;
		CollideArgs collideArgs;
		int collideLanes = 0;
		PackCollideArgs packCollideArgs;
		PrepareArgs prepareArgs;
		int prepareLanes = 0;
		PrepareJointArgs prepareJointArgs;
		int prepareJointLanes = 0;
		FusedArgs fusedArgs;
		StoreArgs storeArgs;
		int storeLanes = 0;
		FinalizeArgs finalizeArgs;
		int finalizeLanes = 0;
		ContinuousArgs continuousArgs;
		int continuousLanes = 0;
		PackFinalizeArgs packFinalizeArgs;

		// host data the batch copies to the device before the next stage it launches
		struct Upload
		{
			void* dst;
			size_t offset;
			size_t bytes;
		};
		std::vector<char> payload;
		std::vector<Upload> uploads;
		bool init = false;
		InitArgs initArgs;
		int initLanes = 0;
	} rec;

	// Host-to-device copy of host data: now on the world's stream, or recorded for the batch.
	void Stage( void* dst, const void* src, size_t bytes )
	{
		if ( bytes == 0 )
		{
			return;
		}
		if ( recording )
		{
			size_t offset = rec.payload.size();
			rec.payload.resize( offset + ( ( bytes + 3 ) & ~size_t( 3 ) ) );
			memcpy( rec.payload.data() + offset, src, bytes );
			rec.uploads.push_back( Record::Upload{ dst, offset, bytes } );
		}
		else
		{
			HIP_CHECK( hipMemcpyAsync( dst, src, bytes, hipMemcpyHostToDevice, stream ) );
		}
	}

	template <typename T>
	void StageVector( DevArray<T>& dst, const std::vector<T>& src )
	{
		dst.Reserve( std::max<size_t>( src.size(), 1 ), false, stream, 0 );
		Stage( dst.ptr, src.data(), src.size() * sizeof( T ) );
	}

	// phase state kept from the launch half to the finish half of a stage
	int collideCount = 0;
	int solveBodyCount = 0;
	bool solveNeedSleepTimes = false;
	PackFinalizeArgs solvePack = {};

	// pinned host mailbox written by the pack kernels
	int* mail = nullptr;
	size_t mailCapacity = 0; // ints

	void EnsureMail( size_t ints )
	{
		if ( ints <= mailCapacity )
		{
			return;
		}
		size_t cap = std::max<size_t>( ints, mailCapacity + mailCapacity / 2 );
		cap = std::max<size_t>( cap, 4096 );
		HIP_CHECK( hipStreamSynchronize( stream ) );
		if ( mail != nullptr )
		{
			HIP_CHECK( hipHostFree( mail ) );
		}
		HIP_CHECK( hipHostMalloc( (void**)&mail, cap * sizeof( int ), hipHostMallocDefault ) );
		mailCapacity = cap;
	}

	hipEvent_t ev[4] = {};
	// XCD solver: barrier words (see SolveXcdKernel) and whether this device can run it
	int* xcdSync = nullptr;
	int xcdUsable = -1; // -1 not checked yet
	// a recorded (batched) world too large for the batch's persistent solver: its solve is launched
	// on its own after the batch's (LaunchDeferredSolve) and waited for in SolveFinish
	bool solveDeferred = false;
	bool deferredLaunched = false;
	StepParams deferredParams = {};
	bool deferredRestitution = false;
	bool deferredSleepTimes = false;
	// recorded after the finalize pack kernel of a launched solve (speculative pair queries wait on it)
	hipEvent_t solvePacked = nullptr;
	bool solvePackedValid = false;
	// the helper thread running this world's solve launches, until JoinLaunch
	StepHelper* launchHelper = nullptr;
	// move-event transforms: the bodies asked for and their gathered transforms
	DevArray<int> transformIds;
	DevArray<b2Transform> transformOut;

	// Per-kernel timing
	bool timing = false;
	struct Pending
	{
		int cls;
		double bytes;
		int ev0;
		int ev1;
	};
	std::vector<Pending> pending;
	std::vector<hipEvent_t> eventPool;
	int eventUsed = 0;
	double statMs[kKernelClassCount] = {};
	double statBytes[kKernelClassCount] = {};
	int statLaunches[kKernelClassCount] = {};

	int NextEvent()
	{
		if ( eventUsed == (int)eventPool.size() )
		{
			hipEvent_t e;
			HIP_CHECK( hipEventCreate( &e ) );
			eventPool.push_back( e );
		}
		return eventUsed++;
	}

	// Launch through this wrapper so the optional timing brackets every kernel.
	template <typename F>
	void Launch( int cls, double bytes, F&& launch )
	{
		// Timed launches hand their events to the dispatch itself (hipExtLaunchKernelGGL), so the
		// events bracket the kernel's execution rather than the queue around it.
		if ( timing )
		{
			int e0 = NextEvent();
			int e1 = NextEvent();
			launch( eventPool[e0], eventPool[e1] );
			HIP_CHECK( hipGetLastError() );
			pending.push_back( Pending{ cls, bytes, e0, e1 } );
		}
		else
		{
			launch( nullptr, nullptr );
			HIP_CHECK( hipGetLastError() );
		}
	}

	// Call after the stream has been synchronized.
	void FlushTiming()
	{
		for ( const Pending& p : pending )
		{
			float ms = 0.0f;
			HIP_CHECK( hipEventElapsedTime( &ms, eventPool[p.ev0], eventPool[p.ev1] ) );
			statMs[p.cls] += ms;
			statBytes[p.cls] += p.bytes;
			statLaunches[p.cls] += 1;
		}
		pending.clear();
		eventUsed = 0;
	}

	void EnsureConstraints( size_t n )
	{
		if ( n <= constraintCapacity )
		{
			return;
		}
		size_t cap = std::max<size_t>( n, constraintCapacity + constraintCapacity / 2 );
		cap = std::max<size_t>( cap, 64 );
		cfields.Free();
		cints.Free();
		cfields.Reserve( cap * kcFieldCount, false, stream, 0 );
		cints.Reserve( cap * 6, false, stream, 0 );
		constraintCapacity = cap;
		soa.f.base = cfields.ptr;
		soa.f.stride = (int)cap;
		soa.bodyA = cints.ptr;
		soa.bodyB = cints.ptr + cap;
		soa.contactId = cints.ptr + 2 * cap;
		soa.pointCount = cints.ptr + 3 * cap;
		soa.slotA = cints.ptr + 4 * cap;
		soa.slotB = cints.ptr + 5 * cap;
	}
};

// Allow the fused solver kernels the whole 160 KB of LDS for dynamic allocation (once per process).
static void EnableLargeLds()
{
	static bool done = false;
	if ( done == false )
	{
		// A step waits on the GPU three times; spinning instead of yielding to the OS while waiting
		// keeps the host's wake-up latency out of each wait. This only takes effect when it is the
		// process's first HIP call; when the runtime is already set up the call fails, and only
		// the error that failure recorded is cleared (b2amd_SpinWaitActive reports the outcome).
		if ( getenv( "B2AMD_SCHEDULE_YIELD" ) == nullptr )
		{
			if ( hipSetDeviceFlags( hipDeviceScheduleSpin ) != hipSuccess )
			{
				(void)hipGetLastError();
			}
		}
		HIP_CHECK( hipFuncSetAttribute( reinterpret_cast<const void*>( SolveFusedKernel ),
										hipFuncAttributeMaxDynamicSharedMemorySize, FusedLdsBytes( kFusedBodyCapacity ) ) );
		HIP_CHECK( hipFuncSetAttribute( reinterpret_cast<const void*>( SolveFusedBatchKernel ),
										hipFuncAttributeMaxDynamicSharedMemorySize, FusedLdsBytes( kFusedBodyCapacity ) ) );
		const void* overflowKernels[4] = { reinterpret_cast<const void*>( OverflowLevelKernel<kStageWarmStart> ),
										   reinterpret_cast<const void*>( OverflowLevelKernel<kStageSolve> ),
										   reinterpret_cast<const void*>( OverflowLevelKernel<kStageRelax> ),
										   reinterpret_cast<const void*>( OverflowLevelKernel<kStageRestitution> ) };
		for ( const void* k : overflowKernels )
		{
			HIP_CHECK( hipFuncSetAttribute( k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)OverflowLdsBytes() ) );
		}
		done = true;
	}
}

DeviceWorld::DeviceWorld()
{
	EnableLargeLds();
	impl = new Impl();
	HIP_CHECK( hipStreamCreateWithFlags( &impl->stream, hipStreamNonBlocking ) );
	HIP_CHECK( hipMalloc( &impl->counters, 16 * sizeof( int ) ) );
	// zero once: from then on the pack kernels clear the counters after reading them
	HIP_CHECK( hipMemset( impl->counters, 0, 16 * sizeof( int ) ) );
	HIP_CHECK( hipHostMalloc( &impl->hostCounters, 16 * sizeof( int ), hipHostMallocDefault ) );
	for ( int i = 0; i < 4; ++i )
	{
		HIP_CHECK( hipEventCreate( &impl->ev[i] ) );
	}
	HIP_CHECK( hipEventCreateWithFlags( &impl->solvePacked, hipEventDisableTiming ) );
	stream = impl->stream;
}

DeviceWorld::~DeviceWorld()
{
	if ( impl == nullptr )
	{
		return;
	}
	JoinLaunch();
	(void)hipStreamSynchronize( impl->stream );
	(void)hipEventDestroy( impl->solvePacked );
	if ( impl->xcdSync != nullptr )
	{
		(void)hipFree( impl->xcdSync );
	}
	impl->sims.Free();
	impl->transformIds.Free();
	impl->transformOut.Free();
	impl->states.Free();
	impl->aux.Free();
	impl->shapes.Free();
	impl->contacts.Free();
	impl->joints.Free();
	impl->inits.Free();
	impl->awakeBodies.Free();
	impl->bodySlot.Free();
	impl->collideList.Free();
	impl->jointList.Free();
	impl->warmStart.Free();
	impl->warmList.Free();
	impl->changes.Free();
	impl->cfields.Free();
	impl->cints.Free();
	impl->enlargedBodies.Free();
	impl->fastBodies.Free();
	impl->ccdTree.Free();
	impl->ccdNoCollide.Free();
	impl->ccdFilterRejects.Free();
	impl->enlargedShapes.Free();
	impl->enlargedShapeAABBs.Free();
	impl->islandSlots.Free();
	impl->islandIds.Free();
	impl->islandBits.Free();
	(void)hipFree( impl->counters );
	(void)hipHostFree( impl->hostCounters );
	if ( impl->mail != nullptr )
	{
		(void)hipHostFree( impl->mail );
	}
	for ( int i = 0; i < 4; ++i )
	{
		(void)hipEventDestroy( impl->ev[i] );
	}
	(void)hipStreamDestroy( impl->stream );
	delete impl;
}

template <typename T>
static void UploadVector( DevArray<T>& dst, const std::vector<T>& src, hipStream_t stream )
{
	dst.Reserve( std::max<size_t>( src.size(), 1 ), false, stream, 0 );
	if ( src.empty() == false )
	{
		HIP_CHECK( hipMemcpyAsync( dst.ptr, src.data(), src.size() * sizeof( T ), hipMemcpyHostToDevice, stream ) );
	}
}

void DeviceWorld::UploadBodies( const World& world )
{
	JoinLaunch();
	int n = (int)world.bodySims.size();
	impl->sims.Reserve( n, false, impl->stream, 0 );
	impl->states.Reserve( n, false, impl->stream, 0 );
	impl->aux.Reserve( n, false, impl->stream, 0 );
	if ( n > 0 )
	{
		HIP_CHECK( hipMemcpyAsync( impl->sims.ptr, world.bodySims.data(), n * sizeof( BodySim ), hipMemcpyHostToDevice, impl->stream ) );
		HIP_CHECK( hipMemcpyAsync( impl->states.ptr, world.bodyStates.data(), n * sizeof( BodyState ), hipMemcpyHostToDevice,
								   impl->stream ) );
		HIP_CHECK( hipMemcpyAsync( impl->aux.ptr, world.bodyAux.data(), n * sizeof( BodyAux ), hipMemcpyHostToDevice, impl->stream ) );
	}
	impl->bodyCount = n;
	HIP_CHECK( hipStreamSynchronize( impl->stream ) );
}

void DeviceWorld::DownloadBodies( World& world )
{
	JoinLaunch();
	int n = impl->bodyCount;
	B2G_CHECK( n <= (int)world.bodySims.size() );
	if ( n > 0 )
	{
		HIP_CHECK( hipMemcpyAsync( world.bodySims.data(), impl->sims.ptr, n * sizeof( BodySim ), hipMemcpyDeviceToHost, impl->stream ) );
		HIP_CHECK( hipMemcpyAsync( world.bodyStates.data(), impl->states.ptr, n * sizeof( BodyState ), hipMemcpyDeviceToHost,
								   impl->stream ) );
		HIP_CHECK( hipMemcpyAsync( world.bodyAux.data(), impl->aux.ptr, n * sizeof( BodyAux ), hipMemcpyDeviceToHost, impl->stream ) );
	}
	HIP_CHECK( hipStreamSynchronize( impl->stream ) );
}

void DeviceWorld::UploadShapes( const World& world )
{
	JoinLaunch();
	int n = (int)world.shapeData.size();
	impl->shapes.Reserve( n, false, impl->stream, 0 );
	if ( n > 0 )
	{
		HIP_CHECK( hipMemcpyAsync( impl->shapes.ptr, world.shapeData.data(), n * sizeof( ShapeDev ), hipMemcpyHostToDevice,
								   impl->stream ) );
	}
	impl->shapeCount = n;
	impl->enlargedShapes.Reserve( std::max<size_t>( n, 1 ), false, impl->stream, 0 );
	impl->enlargedShapeAABBs.Reserve( std::max<size_t>( n, 1 ), false, impl->stream, 0 );

	// the static tree, node for node: continuous collision walks it in the host query order
	{
		const DynamicTree& tree = world.broadPhase.trees[b2_staticBody];
		const std::vector<TreeNode>& nodes = tree.Nodes();
		impl->ccdTreeStaging.resize( nodes.size() );
		for ( size_t i = 0; i < nodes.size(); ++i )
		{
			CcdNode& d = impl->ccdTreeStaging[i];
			d.aabb = nodes[i].aabb;
			d.categoryBits = nodes[i].categoryBits;
			d.leaf = ( nodes[i].flags & kNodeLeaf ) ? 1 : 0;
			d.child1 = d.leaf ? B2_NULL_INDEX : nodes[i].child1;
			d.child2 = d.leaf ? B2_NULL_INDEX : nodes[i].child2;
			d.shapeId = d.leaf ? (int)nodes[i].userData : B2_NULL_INDEX;
		}
		impl->ccdTreeRoot = tree.Root();
		UploadVector( impl->ccdTree, impl->ccdTreeStaging, impl->stream );
	}
	HIP_CHECK( hipStreamSynchronize( impl->stream ) );
}

// The transforms of the listed bodies, in list order (the move events' transforms).
__global__ void GatherTransformsKernel( const int* ids, const BodySim* sims, b2Transform* out, int count )
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if ( i < count )
	{
		out[i] = sims[ids[i]].transform;
	}
}

void DeviceWorld::DownloadTransforms( const int* bodyIds, int count, b2Transform* out )
{
	JoinLaunch();
	if ( count <= 0 )
	{
		return;
	}
	B2G_CHECK( impl->sims.ptr != nullptr );
	hipStream_t s = impl->stream;
	impl->transformIds.Reserve( count, false, s, 0 );
	impl->transformOut.Reserve( count, false, s, 0 );
	HIP_CHECK( hipMemcpyAsync( impl->transformIds.ptr, bodyIds, count * sizeof( int ), hipMemcpyHostToDevice, s ) );
	GatherTransformsKernel<<<Blocks( count ), kBlock, 0, s>>>( impl->transformIds.ptr, impl->sims.ptr, impl->transformOut.ptr, count );
	HIP_CHECK( hipGetLastError() );
	HIP_CHECK( hipMemcpyAsync( out, impl->transformOut.ptr, count * sizeof( b2Transform ), hipMemcpyDeviceToHost, s ) );
	HIP_CHECK( hipStreamSynchronize( s ) );
}

void DeviceWorld::DownloadShapes( World& world )
{
	JoinLaunch();
	int n = impl->shapeCount;
	B2G_CHECK( n <= (int)world.shapeData.size() );
	if ( n > 0 )
	{
		HIP_CHECK( hipMemcpyAsync( world.shapeData.data(), impl->shapes.ptr, n * sizeof( ShapeDev ), hipMemcpyDeviceToHost,
								   impl->stream ) );
	}
	HIP_CHECK( hipStreamSynchronize( impl->stream ) );
}

void DeviceWorld::UploadJoints( const World& world )
{
	int n = (int)world.jointSims.size();
	impl->joints.Reserve( n, false, impl->stream, 0 );
	if ( n > 0 )
	{
		HIP_CHECK( hipMemcpyAsync( impl->joints.ptr, world.jointSims.data(), n * sizeof( JointSim ), hipMemcpyHostToDevice,
								   impl->stream ) );
	}
	impl->jointCount = n;

	// joints that switch off collision between a moving body and a static body
	impl->ccdNoCollideStaging.clear();
	for ( const Joint& joint : world.joints )
	{
		if ( joint.jointId == B2_NULL_INDEX || joint.collideConnected )
		{
			continue;
		}
		int a = joint.edges[0].bodyId;
		int b = joint.edges[1].bodyId;
		bool staticA = world.bodies[a].type == b2_staticBody;
		bool staticB = world.bodies[b].type == b2_staticBody;
		if ( staticA != staticB )
		{
			impl->ccdNoCollideStaging.push_back( staticB ? int2{ a, b } : int2{ b, a } );
		}
	}
	UploadVector( impl->ccdNoCollide, impl->ccdNoCollideStaging, impl->stream );
	HIP_CHECK( hipStreamSynchronize( impl->stream ) );
}

void DeviceWorld::UploadContinuousFilterRejects( const std::vector<uint64_t>& keys )
{
	if ( keys.empty() && impl->ccdFilterRejectCount == 0 )
	{
		return;
	}
	UploadVector( impl->ccdFilterRejects, keys, impl->stream );
	impl->ccdFilterRejectCount = (int)keys.size();
}

// Sensor overlaps (reference src/sensor.c:42-126): lane (sensor, shape) applies the tree query's
// tests (the shape's fat box against the sensor's box, the sensor's mask against the shape's
// category) and the callback's (sensor events on the visitor, another body, the filter, the GJK
// distance below 10 FLT_EPSILON). Every shape is tested instead of walking the trees; the tree walk
// only prunes by those same box and mask tests, so the overlaps are the same.
struct SensorArgs
{
	const int* sensorShapes;
	int sensorCount;
	const ShapeDev* shapes;
	int shapeCount;
	const BodySim* sims;
	IdPair* overlaps;
	int* overlapCount;
	int capacity;
};

__global__ void SensorOverlapKernel( SensorArgs a )
{
	int s = blockIdx.y;
	int j = blockIdx.x * blockDim.x + threadIdx.x;
	if ( s >= a.sensorCount || j >= a.shapeCount )
	{
		return;
	}
	int sensorId = a.sensorShapes[s];
	if ( j == sensorId )
	{
		return;
	}
	const ShapeDev& other = a.shapes[j];
	const uint32_t need = kShapeHasProxy | kShapeEnableSensorEvents;
	if ( ( other.flags & need ) != need )
	{
		return;
	}
	const ShapeDev& sensor = a.shapes[sensorId];
	if ( b2AABB_Overlaps( other.fatAABB, sensor.aabb ) == false || ( other.filter.categoryBits & sensor.filter.maskBits ) == 0 )
	{
		return;
	}
	if ( other.bodyId == sensor.bodyId || ShouldShapesCollide( sensor.filter, other.filter ) == false )
	{
		return;
	}
	b2DistanceInput input;
	input.proxyA = ShapeProxy( sensor );
	input.proxyB = ShapeProxy( other );
	input.transform = b2InvMulWorldTransforms( a.sims[sensor.bodyId].transform, a.sims[other.bodyId].transform );
	input.useRadii = true;
	b2SimplexCache cache = {};
	b2DistanceOutput output = ShapeDistance( input, &cache );
	if ( output.distance < 10.0f * FLT_EPSILON )
	{
		int slot = atomicAdd( a.overlapCount, 1 );
		if ( slot < a.capacity )
		{
			a.overlaps[slot] = IdPair{ s, j };
		}
	}
}

void DeviceWorld::SensorOverlaps( const std::vector<int>& sensorShapes, std::vector<IdPair>* overlaps )
{
	overlaps->clear();
	int n = (int)sensorShapes.size();
	if ( n == 0 || impl->shapeCount == 0 )
	{
		return;
	}
	hipStream_t s = impl->stream;
	UploadVector( impl->sensorShapes, sensorShapes, s );
	int* count = impl->counters + kSensorOverlapCountSlot;
	for ( int attempt = 0; attempt < 2; ++attempt )
	{
		if ( impl->sensorOverlapCapacity == 0 )
		{
			impl->sensorOverlapCapacity = 4096;
		}
		impl->sensorOverlaps.Reserve( impl->sensorOverlapCapacity, false, s, 0 );
		SensorArgs a;
		a.sensorShapes = impl->sensorShapes.ptr;
		a.sensorCount = n;
		a.shapes = impl->shapes.ptr;
		a.shapeCount = impl->shapeCount;
		a.sims = impl->sims.ptr;
		a.overlaps = impl->sensorOverlaps.ptr;
		a.overlapCount = count;
		a.capacity = impl->sensorOverlapCapacity;
		HIP_CHECK( hipMemsetAsync( count, 0, sizeof( int ), s ) );
		SensorOverlapKernel<<<dim3( Blocks( impl->shapeCount ), n ), kBlock, 0, s>>>( a );
		HIP_CHECK( hipGetLastError() );
		HIP_CHECK( hipMemcpyAsync( impl->hostCounters + kSensorOverlapCountSlot, count, sizeof( int ), hipMemcpyDeviceToHost, s ) );
		HIP_CHECK( hipStreamSynchronize( s ) );
		int found = impl->hostCounters[kSensorOverlapCountSlot];
		if ( found > impl->sensorOverlapCapacity )
		{
			// too many to keep: grow and run again
			impl->sensorOverlapCapacity = 2 * found;
			continue;
		}
		overlaps->resize( found );
		if ( found > 0 )
		{
			HIP_CHECK( hipMemcpyAsync( overlaps->data(), impl->sensorOverlaps.ptr, found * sizeof( IdPair ), hipMemcpyDeviceToHost, s ) );
			HIP_CHECK( hipStreamSynchronize( s ) );
		}
		return;
	}
	Fatal( "sensor overlaps kept growing between two identical queries" );
}

void DeviceWorld::ContinuousSensorHits( std::vector<IdPair>* hits )
{
	hits->clear();
	if ( impl->sensorHitCapacity == 0 )
	{
		return;
	}
	hipStream_t s = impl->stream;
	int* count = impl->counters + kSensorHitCountSlot;
	HIP_CHECK( hipMemcpyAsync( impl->hostCounters + kSensorHitCountSlot, count, sizeof( int ), hipMemcpyDeviceToHost, s ) );
	HIP_CHECK( hipMemsetAsync( count, 0, sizeof( int ), s ) );
	HIP_CHECK( hipStreamSynchronize( s ) );
	int found = impl->hostCounters[kSensorHitCountSlot];
	if ( found > impl->sensorHitCapacity )
	{
		Fatal( "continuous sensor hits %d exceed their capacity %d", found, impl->sensorHitCapacity );
	}
	hits->resize( found );
	if ( found > 0 )
	{
		HIP_CHECK( hipMemcpyAsync( hits->data(), impl->sensorHits.ptr, found * sizeof( IdPair ), hipMemcpyDeviceToHost, s ) );
		HIP_CHECK( hipStreamSynchronize( s ) );
	}
}

void DeviceWorld::DownloadJoints( World& world )
{
	int n = impl->jointCount;
	B2G_CHECK( n <= (int)world.jointSims.size() );
	if ( n > 0 )
	{
		HIP_CHECK( hipMemcpyAsync( world.jointSims.data(), impl->joints.ptr, n * sizeof( JointSim ), hipMemcpyDeviceToHost,
								   impl->stream ) );
	}
	HIP_CHECK( hipStreamSynchronize( impl->stream ) );
}

void DeviceWorld::UploadContactInits( const std::vector<ContactInit>& inits, int contactCapacity )
{
	impl->contacts.Reserve( contactCapacity, true, impl->stream, impl->contactCapacityUsed );
	impl->contactCapacityUsed = std::max( impl->contactCapacityUsed, contactCapacity );

	int n = (int)inits.size();
	if ( n == 0 )
	{
		return;
	}
	B2G_CHECK( impl->shapeCount > 0 );

	impl->inits.Reserve( n, false, impl->stream, 0 );
	impl->Stage( impl->inits.ptr, inits.data(), n * sizeof( ContactInit ) );
	InitArgs ia = { impl->contacts.ptr, impl->inits.ptr, n };
	if ( impl->recording )
	{
		impl->rec.init = true;
		impl->rec.initArgs = ia;
		impl->rec.initLanes = n;
		return;
	}
	LaneKernel<InitContactsOp><<<Blocks( n ), kBlock, 0, impl->stream>>>( ia );
	HIP_CHECK( hipGetLastError() );
	HIP_CHECK( hipStreamSynchronize( impl->stream ) );
}

__global__ void MarkMixStaleKernel( ContactSim* contacts, int count )
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if ( i < count )
	{
		contacts[i].simFlags |= kSimMixStale;
	}
}

void DeviceWorld::MarkContactMixingStale()
{
	int n = impl->contactCapacityUsed;
	if ( n == 0 )
	{
		return;
	}
	// on the world's own stream and waited for, so a batch's later launches see it too
	MarkMixStaleKernel<<<Blocks( n ), kBlock, 0, impl->stream>>>( impl->contacts.ptr, n );
	HIP_CHECK( hipGetLastError() );
	HIP_CHECK( hipStreamSynchronize( impl->stream ) );
}

void DeviceWorld::UploadContactMaterials( const std::vector<ContactMaterial>& list )
{
	static_assert( offsetof( ContactSim, restitution ) == offsetof( ContactSim, friction ) + sizeof( float ), "adjacent" );
	for ( const ContactMaterial& m : list )
	{
		float values[2] = { m.friction, m.restitution };
		impl->Stage( &impl->contacts.ptr[m.contactId].friction, values, sizeof( values ) );
	}
}

void DeviceWorld::DownloadContacts( int count, ContactSim* out )
{
	count = std::min( count, impl->contactCapacityUsed );
	if ( count <= 0 )
	{
		return;
	}
	HIP_CHECK( hipMemcpyAsync( out, impl->contacts.ptr, count * sizeof( ContactSim ), hipMemcpyDeviceToHost, impl->stream ) );
	HIP_CHECK( hipStreamSynchronize( impl->stream ) );
}

void DeviceWorld::UploadContacts( const ContactSim* sims, int count )
{
	impl->contacts.Reserve( std::max( count, 1 ), false, impl->stream, 0 );
	impl->contactCapacityUsed = std::max( impl->contactCapacityUsed, count );
	if ( count > 0 )
	{
		HIP_CHECK( hipMemcpyAsync( impl->contacts.ptr, sims, count * sizeof( ContactSim ), hipMemcpyHostToDevice, impl->stream ) );
	}
	HIP_CHECK( hipStreamSynchronize( impl->stream ) );
}

void DeviceWorld::DownloadContact( int contactId, ContactSim* out )
{
	HIP_CHECK( hipMemcpyAsync( out, impl->contacts.ptr + contactId, sizeof( ContactSim ), hipMemcpyDeviceToHost, impl->stream ) );
	HIP_CHECK( hipStreamSynchronize( impl->stream ) );
}


void DeviceWorld::UploadWarmLists( const SolverLists& lists )
{
	impl->lists.warmStart = lists.warmStart;
	impl->lists.warmList = lists.warmList;
	impl->StageVector( impl->warmStart, lists.warmStart );
	impl->StageVector( impl->warmList, lists.warmList );
}

void DeviceWorld::UploadLists( const SolverLists& lists )
{
	impl->lists = lists;
	impl->StageVector( impl->awakeBodies, lists.awakeBodies );
	{
		int slots = impl->bodyCount;
		for ( int id : lists.awakeBodies )
		{
			slots = std::max( slots, id + 1 );
		}
		impl->slotStaging.assign( std::max( slots, 1 ), kStaticSlot );
		for ( size_t k = 0; k < lists.awakeBodies.size(); ++k )
		{
			impl->slotStaging[lists.awakeBodies[k]] = (int)std::min<size_t>( k + 1, kFusedBodyCapacity );
		}
		impl->StageVector( impl->bodySlot, impl->slotStaging );
	}
	impl->StageVector( impl->collideList, lists.collideContacts );
	impl->StageVector( impl->jointList, lists.jointList );
	impl->StageVector( impl->warmStart, lists.warmStart );
	impl->StageVector( impl->warmList, lists.warmList );
	impl->StageVector( impl->overflowPlan, lists.overflowPlan );

	size_t nc = lists.constraintContacts.size();
	impl->EnsureConstraints( std::max<size_t>( nc, 1 ) );
	impl->Stage( impl->soa.contactId, lists.constraintContacts.data(), nc * sizeof( int ) );

	size_t nb = lists.awakeBodies.size();
	impl->enlargedBodies.Reserve( std::max<size_t>( nb, 1 ), false, impl->stream, 0 );
	impl->fastBodies.Reserve( std::max<size_t>( nb, 1 ), false, impl->stream, 0 );

	// at most two entries per contact (a touch change and a mixing refresh), then the contacts
	// deferred to the host's pre-solve callback
	impl->changes.Reserve( std::max<size_t>( 3 * lists.collideContacts.size(), 1 ), false, impl->stream, 0 );
	if ( impl->recording == false )
	{
		HIP_CHECK( hipStreamSynchronize( impl->stream ) );
	}
}

void DeviceWorld::CollideLaunch( const StepParams& params )
{
	int n = (int)impl->lists.collideContacts.size();
	impl->collideCount = n;
	if ( n == 0 )
	{
		return;
	}

	hipStream_t s = impl->stream;
	impl->EnsureMail( kMailHeader + 3 * (size_t)n );

	CollideArgs a = {};
	a.list = impl->collideList.ptr;
	a.count = n;
	a.contacts = impl->contacts.ptr;
	a.shapes = impl->shapes.ptr;
	a.sims = impl->sims.ptr;
	a.aux = impl->aux.ptr;
	a.changes = impl->changes.ptr;
	a.counters = impl->counters;
	a.recycleDistance = params.contactRecycleDistance;
	a.customMixing = params.customMixing;
	a.preSolve = params.preSolve;
	a.deferred = impl->changes.ptr + 2 * n;

	PackCollideArgs pc = { impl->counters, impl->changes.ptr, impl->mail, 2 * n, 0 };

	if ( impl->recording )
	{
		impl->rec.collide = true;
		impl->rec.collideArgs = a;
		impl->rec.collideLanes = n;
		impl->rec.packCollide = true;
		impl->rec.packCollideArgs = pc;
		return;
	}

	HIP_CHECK( hipEventRecord( impl->ev[0], s ) );
	// Bytes are fixed up after the launch once the recycled count is known. One launch with its
	// own pack (B2AMD_NO_FUSED_FINISH: the narrowphase and the pack as two launches).
	static const bool fused = getenv( "B2AMD_NO_FUSED_FINISH" ) == nullptr;
	if ( fused )
	{
		int* done = impl->counters + kCollideDoneSlot;
		impl->Launch( kkCollide, 0.0, [&]( hipEvent_t e0, hipEvent_t e1 ) {
			hipExtLaunchKernelGGL( CollideAndPackKernel, dim3( Blocks( n ) ), dim3( kBlock ), 0, s, e0, e1, 0, a, pc, done );
		} );
		HIP_CHECK( hipEventRecord( impl->ev[1], s ) );
		return;
	}
	impl->Launch( kkCollide, 0.0, [&]( hipEvent_t e0, hipEvent_t e1 ) { hipExtLaunchKernelGGL( LaneKernel<CollideOp>, dim3( Blocks( n ) ), dim3( kBlock ), 0, s, e0, e1, 0, a ); } );
	HIP_CHECK( hipEventRecord( impl->ev[1], s ) );
	PackCollideKernel<<<1, kPackThreads, 0, s>>>( pc );
	HIP_CHECK( hipGetLastError() );
}

void DeviceWorld::CollideFinish( CollideResult* result )
{
	result->changes.clear();
	result->deferred.clear();
	result->recycledCount = 0;

	int n = impl->collideCount;
	if ( n == 0 )
	{
		return;
	}

	if ( impl->recording == false )
	{
		HIP_CHECK( hipStreamSynchronize( impl->stream ) );
	}

	int changeCount = impl->mail[0];
	result->recycledCount = impl->mail[1];
	if ( impl->timing && impl->pending.empty() == false )
	{
		int recycled = result->recycledCount;
		impl->pending.back().bytes = kBytesCollideFull * ( n - recycled ) + kBytesCollideRecycled * recycled;
	}
	impl->FlushTiming();

	if ( changeCount > 0 )
	{
		result->changes.assign( impl->mail + kMailHeader, impl->mail + kMailHeader + changeCount );
		std::sort( result->changes.begin(), result->changes.end() );
	}
	int deferredCount = impl->mail[2];
	if ( deferredCount < 0 || deferredCount > n )
	{
		Fatal( "pre-solve contact count %d exceeds the collided contact count %d", deferredCount, n );
	}
	result->deferred.assign( impl->mail + kMailHeader + 2 * n, impl->mail + kMailHeader + 2 * n + deferredCount );
	std::sort( result->deferred.begin(), result->deferred.end() );

	if ( impl->recording == false )
	{
		HIP_CHECK( hipEventElapsedTime( &lastCollideMs, impl->ev[0], impl->ev[1] ) );
	}
}

// Host sink of the narrowphase: change records into a vector and the world's pre-solve callback
// on the new manifold, before anchors move to the centers of mass (reference src/contact.c:573-583).
struct HostCollideSink
{
	std::vector<int>* changes;
	const World* world;

	void Change( int code ) const
	{
		changes->push_back( code );
	}
	void Recycled() const
	{
	}
	bool CheckRecycle() const
	{
		// the device found these contacts not recyclable before deferring them
		return false;
	}
	bool Defer( int, uint32_t ) const
	{
		return false;
	}
	void PreSolve( const ContactSim& cs, b2Manifold* manifold ) const
	{
		world->preSolveFcn( world->ShapeIdOf( cs.shapeIdA ), world->ShapeIdOf( cs.shapeIdB ), manifold, world->preSolveContext );
	}
};

void DeviceWorld::CollideOnHost( World& world, const std::vector<int>& contactIds, const StepParams& params,
								 std::vector<int>* changes )
{
	CollideArgs a = {};
	a.list = contactIds.data();
	a.count = (int)contactIds.size();
	a.contacts = world.contactSims.data();
	a.shapes = world.shapeData.data();
	a.sims = world.bodySims.data();
	a.aux = world.bodyAux.data();
	a.recycleDistance = params.contactRecycleDistance;
	a.customMixing = params.customMixing;
	a.preSolve = 1;
	HostCollideSink sink{ changes, &world };
	for ( int k = 0; k < a.count; ++k )
	{
		CollideBody( a, k, sink );
	}
	for ( int contactId : contactIds )
	{
		HIP_CHECK( hipMemcpyAsync( impl->contacts.ptr + contactId, world.contactSims.data() + contactId, sizeof( ContactSim ),
								   hipMemcpyHostToDevice, impl->stream ) );
	}
	HIP_CHECK( hipStreamSynchronize( impl->stream ) );
}

// Continuous collision of one fast body on the host (reference src/solver.c:181-550): the callback
// tests of b2ContinuousQueryCallback in its order, including the two the device kernel cannot run
// (the custom filter and the pre-continuous callback, called here as the reference calls them),
// and the bullet branch that also sweeps the kinematic and dynamic trees, skipping other bullets.
// Reads and writes the host copies of bodies and shapes.
void DeviceWorld::ContinuousOnHost( World& world, int bodyId, std::vector<IdPair>* sensorHits )
{
	BodySim* fastSim = &world.bodySims[bodyId];
	const bool isBullet = ( fastSim->flags & kIsBullet ) != 0;
	b2Vec2 base = fastSim->center0;
	b2Sweep sweep = RelativeSweep( *fastSim, base );

	b2Transform xf1;
	xf1.q = sweep.q1;
	xf1.p = b2Sub( sweep.c1, b2RotateVector( sweep.q1, sweep.localCenter ) );
	b2Transform xf2;
	xf2.q = sweep.q2;
	xf2.p = b2Sub( sweep.c2, b2RotateVector( sweep.q2, sweep.localCenter ) );

	float fraction = 1.0f;
	IdPair hits[kMaxContinuousSensorHits];
	float hitFractions[kMaxContinuousSensorHits];
	int hitCount = 0;
	Body* fastBody = &world.bodies[bodyId];

	for ( int shapeId = fastSim->headShapeId; shapeId != B2_NULL_INDEX; shapeId = world.shapeData[shapeId].nextShapeId )
	{
		ShapeDev* fastShape = &world.shapeData[shapeId];
		const Shape& fastHost = world.shapes[shapeId];
		b2Vec2 centroid1 = b2TransformPoint( xf1, fastShape->localCentroid );
		b2Vec2 centroid2 = b2TransformPoint( xf2, fastShape->localCentroid );

		b2AABB box1 = fastShape->aabb;
		b2AABB box2 = OffsetAABB( FatShapeAABB( *fastShape, xf2, 0.0f ), base );
		fastShape->aabb = box2;
		if ( fastShape->flags & kShapeIsSensor )
		{
			continue;
		}
		b2AABB swept = b2AABB_Union( box1, box2 );

		auto visit = [&]( int, uint64_t userData ) -> bool {
			int otherId = (int)userData;
			if ( otherId == shapeId )
			{
				return true;
			}
			const ShapeDev& other = world.shapeData[otherId];
			const Shape& otherHost = world.shapes[otherId];
			if ( other.bodyId == fastShape->bodyId )
			{
				return true;
			}
			const bool otherSensor = ( other.flags & kShapeIsSensor ) != 0;
			if ( otherSensor && ( other.flags & fastShape->flags & kShapeEnableSensorEvents ) == 0 )
			{
				return true;
			}
			if ( ShouldShapesCollide( fastShape->filter, other.filter ) == false )
			{
				return true;
			}
			const BodySim& otherSim = world.bodySims[other.bodyId];
			if ( otherSim.flags & kIsBullet )
			{
				return true;
			}
			if ( world.ShouldBodiesCollide( fastBody, &world.bodies[other.bodyId] ) == false )
			{
				return true;
			}
			if ( ( otherHost.enableCustomFiltering || fastHost.enableCustomFiltering ) && world.customFilterFcn != nullptr &&
				 world.customFilterFcn( world.ShapeIdOf( otherId ), world.ShapeIdOf( shapeId ), world.customFilterContext ) == false )
			{
				return true;
			}

			if ( other.type == b2_chainSegmentShape )
			{
				b2Transform rel = { b2SubPos( otherSim.transform.p, base ), otherSim.transform.q };
				b2Vec2 p1 = b2TransformPoint( rel, other.chainSegment.segment.point1 );
				b2Vec2 p2 = b2TransformPoint( rel, other.chainSegment.segment.point2 );
				float length;
				b2Vec2 e = b2GetLengthAndNormalize( &length, b2Sub( p2, p1 ) );
				if ( length > kLinearSlop )
				{
					float separation1 = b2Cross( b2Sub( centroid1, p1 ), e );
					float separation2 = b2Cross( b2Sub( centroid2, p1 ), e );
					float coreDistance = kCoreFraction * fastSim->minExtent;
					if ( separation1 < 0.0f || ( separation1 - separation2 < coreDistance && separation2 > coreDistance ) )
					{
						return true;
					}
				}
			}

			b2TOIInput input;
			input.proxyA = ShapeProxy( other );
			input.proxyB = ShapeProxy( *fastShape );
			input.sweepA = RelativeSweep( otherSim, base );
			input.sweepB = sweep;
			input.maxFraction = fraction;

			b2TOIOutput output = TimeOfImpact( input );
			if ( otherSensor )
			{
				if ( output.fraction <= fraction && hitCount < kMaxContinuousSensorHits )
				{
					hits[hitCount] = IdPair{ otherId, shapeId };
					hitFractions[hitCount] = output.fraction;
					hitCount += 1;
				}
				return true;
			}

			float hitFraction = fraction;
			bool didHit = false;
			if ( 0.0f < output.fraction && output.fraction < fraction )
			{
				hitFraction = output.fraction;
				didHit = true;
			}
			else if ( 0.0f == output.fraction )
			{
				b2Vec2 centroid = ShapeCentroidDev( *fastShape );
				float radius = kCoreFraction * ShapeMinExtent( *fastShape, B2_HUGE );
				input.proxyB = MakeProxy( &centroid, 1, radius );
				output = TimeOfImpact( input );
				if ( 0.0f < output.fraction && output.fraction < fraction )
				{
					hitFraction = output.fraction;
					didHit = true;
				}
			}

			if ( didHit && ( otherHost.enablePreSolveEvents || fastHost.enablePreSolveEvents ) && world.preContinuousFcn != nullptr )
			{
				// the time of impact is in the base frame: the callback gets the world point
				b2Vec2 worldPoint = b2OffsetPos( base, output.point );
				didHit = world.preContinuousFcn( world.ShapeIdOf( otherId ), world.ShapeIdOf( shapeId ), worldPoint, output.normal,
												 world.preSolveContext );
			}

			if ( didHit )
			{
				fastSim->flags |= kHadTimeOfImpact;
				fraction = hitFraction;
			}
			return true;
		};

		world.broadPhase.trees[b2_staticBody].Query( swept, B2_DEFAULT_MASK_BITS, visit );
		if ( isBullet )
		{
			world.broadPhase.trees[b2_kinematicBody].Query( swept, B2_DEFAULT_MASK_BITS, visit );
			world.broadPhase.trees[b2_dynamicBody].Query( swept, B2_DEFAULT_MASK_BITS, visit );
		}
	}

	auto enlarge = [&]( ShapeDev* shape ) {
		if ( b2AABB_Contains( shape->fatAABB, shape->aabb ) == false )
		{
			float margin = shape->aabbMargin;
			b2AABB fat;
			fat.lowerBound.x = shape->aabb.lowerBound.x - margin;
			fat.lowerBound.y = shape->aabb.lowerBound.y - margin;
			fat.upperBound.x = shape->aabb.upperBound.x + margin;
			fat.upperBound.y = shape->aabb.upperBound.y + margin;
			shape->fatAABB = fat;
			shape->enlargedAABB = 1;
			fastSim->flags |= kEnlargeBounds;
		}
	};

	if ( fraction < 1.0f )
	{
		b2Rot q = b2NLerp( sweep.q1, sweep.q2, fraction );
		b2Vec2 c = b2Lerp( sweep.c1, sweep.c2, fraction );
		b2Vec2 origin = b2Sub( c, b2RotateVector( q, sweep.localCenter ) );

		fastSim->transform.q = q;
		fastSim->transform.p = b2OffsetPos( base, origin );
		fastSim->center = b2OffsetPos( base, c );
		fastSim->rotation0 = q;
		fastSim->center0 = fastSim->center;

		b2Transform transform = fastSim->transform;
		for ( int shapeId = fastSim->headShapeId; shapeId != B2_NULL_INDEX; shapeId = world.shapeData[shapeId].nextShapeId )
		{
			ShapeDev* shape = &world.shapeData[shapeId];
			shape->aabb = FatShapeAABB( *shape, transform, kSpeculativeDistance );
			enlarge( shape );
		}
	}
	else
	{
		fastSim->rotation0 = fastSim->transform.q;
		fastSim->center0 = fastSim->center;
		for ( int shapeId = fastSim->headShapeId; shapeId != B2_NULL_INDEX; shapeId = world.shapeData[shapeId].nextShapeId )
		{
			enlarge( &world.shapeData[shapeId] );
		}
	}

	for ( int i = 0; i < hitCount; ++i )
	{
		if ( hitFractions[i] < fraction )
		{
			sensorHits->push_back( hits[i] );
		}
	}
}

void DeviceWorld::UploadBodyShapes( const World& world, const std::vector<int>& bodyIds )
{
	for ( int bodyId : bodyIds )
	{
		HIP_CHECK( hipMemcpyAsync( impl->sims.ptr + bodyId, world.bodySims.data() + bodyId, sizeof( BodySim ), hipMemcpyHostToDevice,
								   impl->stream ) );
		for ( int shapeId = world.bodySims[bodyId].headShapeId; shapeId != B2_NULL_INDEX; shapeId = world.shapeData[shapeId].nextShapeId )
		{
			HIP_CHECK( hipMemcpyAsync( impl->shapes.ptr + shapeId, world.shapeData.data() + shapeId, sizeof( ShapeDev ),
									   hipMemcpyHostToDevice, impl->stream ) );
		}
	}
	HIP_CHECK( hipStreamSynchronize( impl->stream ) );
}

void DeviceWorld::Collide( const StepParams& params, CollideResult* result )
{
	CollideLaunch( params );
	CollideFinish( result );
}

// Warm-start form of a solver path. B2AMD_WARM_BY_BODY: unset = defaults (body by body for the
// color kernels, color by color in the fused solver), 0 = color by color everywhere, 1 = body by
// body everywhere. Results are identical; only the time differs.
static bool WarmByBody( bool fusedPath )
{
	static const int mode = []() {
		const char* v = getenv( "B2AMD_WARM_BY_BODY" );
		return v != nullptr ? atoi( v ) : -1;
	}();
	if ( mode >= 0 )
	{
		return mode != 0;
	}
	return fusedPath == false;
}

// Batched persistent solver configuration by world-size bucket (the offline table, overridable).
static const BatchTuning kBatchTuningTable[kBatchBuckets] = B2AMD_BATCH_TUNING_TABLE;
static std::atomic<int> s_bucketThreads[kBatchBuckets];	 // 0: the table's
static std::atomic<int> s_bucketWarmByBody[kBatchBuckets]; // -1: the table's
static std::atomic<bool> s_bucketOverridesInit{ false };
// Workgroup size forced for every bucket by b2amd_SetBatchSolverThreads (0: per bucket).
static std::atomic<int> s_forcedBatchThreads{ 0 };

static bool ValidBatchThreads( int threads )
{
	return threads == 64 || threads == 128 || threads == 256 || threads == 512 || threads == 1024;
}

static void InitBucketOverrides()
{
	if ( s_bucketOverridesInit.load() == false )
	{
		for ( int k = 0; k < kBatchBuckets; ++k )
		{
			s_bucketThreads[k].store( 0 );
			s_bucketWarmByBody[k].store( -1 );
		}
		const char* v = getenv( "B2AMD_BATCH_THREADS" );
		if ( v != nullptr && ValidBatchThreads( atoi( v ) ) )
		{
			s_forcedBatchThreads.store( atoi( v ) );
		}
		s_bucketOverridesInit.store( true );
	}
}

int BatchBucketOf( int bodyCount )
{
	for ( int k = 0; k < kBatchBuckets - 1; ++k )
	{
		if ( bodyCount <= kBatchTuningTable[k].maxBodies )
		{
			return k;
		}
	}
	return kBatchBuckets - 1;
}

BatchTuning BatchTuningTable( int bucket )
{
	return kBatchTuningTable[bucket];
}

BatchTuning BatchTuningFor( int bucket )
{
	InitBucketOverrides();
	BatchTuning t = kBatchTuningTable[bucket];
	int threads = s_bucketThreads[bucket].load();
	int warm = s_bucketWarmByBody[bucket].load();
	int forced = s_forcedBatchThreads.load();
	t.threads = forced != 0 ? forced : ( threads != 0 ? threads : t.threads );
	t.warmByBody = warm >= 0 ? warm : t.warmByBody;
	return t;
}

bool SetBatchTuning( int bucket, int threads, int warmByBody )
{
	if ( bucket < 0 || bucket >= kBatchBuckets || ( threads != 0 && ValidBatchThreads( threads ) == false ) ||
		 ( threads != 0 && warmByBody != 0 && warmByBody != 1 ) )
	{
		return false;
	}
	InitBucketOverrides();
	s_bucketThreads[bucket].store( threads );
	s_bucketWarmByBody[bucket].store( threads == 0 ? -1 : warmByBody );
	return true;
}

bool SetBatchSolverThreads( int threads )
{
	if ( threads != 0 && ValidBatchThreads( threads ) == false )
	{
		return false;
	}
	InitBucketOverrides();
	s_forcedBatchThreads.store( threads );
	return true;
}

// The warm-start form of a batched world with bodyCount awake bodies: B2AMD_WARM_BY_BODY when set,
// else its bucket's.
static bool BatchWarmByBody( int bodyCount )
{
	const char* v = getenv( "B2AMD_WARM_BY_BODY" );
	if ( v != nullptr )
	{
		return atoi( v ) != 0;
	}
	return BatchTuningFor( BatchBucketOf( bodyCount ) ).warmByBody != 0;
}

// Whether the current device can run the XCD solver: an MI355X (8 XCDs of 32 CUs) with
// cooperative launches and room for one of its workgroups on every CU.
static bool XcdSolverUsable()
{
	if ( getenv( "B2AMD_NO_XCD_SOLVER" ) != nullptr )
	{
		return false;
	}
	int device = 0;
	hipDeviceProp_t prop;
	if ( hipGetDevice( &device ) != hipSuccess || hipGetDeviceProperties( &prop, device ) != hipSuccess )
	{
		(void)hipGetLastError();
		return false;
	}
	int perCu = 0;
	if ( hipOccupancyMaxActiveBlocksPerMultiprocessor( &perCu, SolveXcdKernel, kXcdThreads, 0 ) != hipSuccess )
	{
		(void)hipGetLastError();
		return false;
	}
	return prop.cooperativeLaunch != 0 && prop.multiProcessorCount == kXcdGrid && perCu >= 1 &&
		   strncmp( prop.gcnArchName, "gfx950", 6 ) == 0;
}

void DeviceWorld::ArmLaunchHelper()
{
	StepHelper* helper = impl->recording ? nullptr : SharedStepHelper();
	if ( helper != nullptr )
	{
		helper->Arm();
	}
}

void DeviceWorld::SolveLaunchAsync( const StepParams& params, bool anyRestitution, bool needSleepTimes )
{
	StepHelper* helper = impl->recording ? nullptr : SharedStepHelper();
	if ( helper != nullptr )
	{
		int device = 0;
		HIP_CHECK( hipGetDevice( &device ) );
		StepParams p = params;
		if ( helper->TryPost( [this, p, anyRestitution, needSleepTimes, device]() {
				 HIP_CHECK( hipSetDevice( device ) );
				 SolveLaunch( p, anyRestitution, needSleepTimes );
			 } ) )
		{
			impl->launchHelper = helper;
			return;
		}
	}
	SolveLaunch( params, anyRestitution, needSleepTimes );
}

void DeviceWorld::JoinLaunch() const
{
	if ( impl->launchHelper != nullptr )
	{
		impl->launchHelper->Wait();
		impl->launchHelper = nullptr;
	}
}

void DeviceWorld::SolveLaunch( const StepParams& params, bool anyRestitution, bool needSleepTimes )
{
	const SolverLists& L = impl->lists;
	hipStream_t s = impl->stream;
	int nb = (int)L.awakeBodies.size();
	int nc = (int)L.constraintContacts.size();
	int nj = (int)L.jointList.size();
	const bool rec = impl->recording;

	if ( rec && nb > 0 && impl->tape == nullptr && ( nb > kFusedBodyCapacity || L.activeColorCount > kMaxColorSlots ) )
	{
		// beyond the batch's persistent solver: this world's solve runs alone after the batch's
		impl->solveDeferred = true;
		impl->deferredParams = params;
		impl->deferredRestitution = anyRestitution;
		impl->deferredSleepTimes = needSleepTimes;
		return;
	}

	if ( rec == false )
	{
		HIP_CHECK( hipEventRecord( impl->ev[2], s ) );
	}

	if ( nc > 0 )
	{
		PrepareArgs pa;
		pa.c = impl->soa;
		pa.count = nc;
		pa.contacts = impl->contacts.ptr;
		pa.sims = impl->sims.ptr;
		pa.states = impl->states.ptr;
		pa.aux = impl->aux.ptr;
		pa.bodySlot = impl->bodySlot.ptr;
		pa.p = params;
		if ( rec )
		{
			impl->rec.prepare = true;
			impl->rec.prepareArgs = pa;
			impl->rec.prepareLanes = nc;
		}
		else
		{
			impl->Launch( kkPrepareContacts, kBytesPrepare * nc,
						  [&]( hipEvent_t e0, hipEvent_t e1 ) { hipExtLaunchKernelGGL( LaneKernel<PrepareContactsOp>, dim3( ( nc + kWarmBlock - 1 ) / kWarmBlock ), dim3( kWarmBlock ), 0, s, e0, e1, 0, pa ); } );
		}
	}

	if ( nj > 0 )
	{
		PrepareJointArgs ja;
		ja.joints = impl->joints.ptr;
		ja.list = impl->jointList.ptr;
		ja.count = nj;
		ja.sims = impl->sims.ptr;
		ja.aux = impl->aux.ptr;
		ja.bodySlot = impl->bodySlot.ptr;
		ja.p = params;
		if ( rec )
		{
			impl->rec.prepareJoints = true;
			impl->rec.prepareJointArgs = ja;
			impl->rec.prepareJointLanes = nj;
		}
		else
		{
			impl->Launch( kkPrepareJoints, kBytesJoint * nj,
						  [&]( hipEvent_t e0, hipEvent_t e1 ) { hipExtLaunchKernelGGL( LaneKernel<PrepareJointsOp>, dim3( Blocks( nj ) ), dim3( kBlock ), 0, s, e0, e1, 0, ja ); } );
		}
	}

	BodyArgs ba;
	ba.awake = impl->awakeBodies.ptr;
	ba.count = nb;
	ba.sims = impl->sims.ptr;
	ba.states = impl->states.ptr;
	ba.aux = impl->aux.ptr;
	ba.p = params;

	ColorArgs ca;
	ca.c = impl->soa;
	ca.joints = impl->joints.ptr;
	ca.jointList = impl->jointList.ptr;
	ca.states = impl->states.ptr;
	ca.p = params;

	auto stageClass = []( int stage ) {
		switch ( stage )
		{
			case kStageWarmStart:
				return (int)kkWarmStart;
			case kStageSolve:
				return (int)kkSolve;
			case kStageRelax:
				return (int)kkRelax;
			default:
				return (int)kkRestitution;
		}
	};
	auto stageBytes = []( int stage ) {
		switch ( stage )
		{
			case kStageWarmStart:
				return kBytesWarmStart;
			case kStageSolve:
				return kBytesSolve;
			case kStageRelax:
				return kBytesRelax;
			default:
				return kBytesRestitution;
		}
	};

	const int cb = impl->colorBlock;
	auto runStage = [&]( int stage ) {
		if ( L.overflowContactCount > 0 || L.overflowJointCount > 0 )
		{
			ca.stage = stage;
			ca.jointStart = L.overflowJointStart;
			ca.jointCount = L.overflowJointCount;
			ca.contactStart = L.overflowContactStart;
			ca.contactCount = L.overflowContactCount;
			double bytes = stageBytes( stage ) * ca.contactCount + kBytesJoint * ca.jointCount;
			const int* plan = impl->overflowPlan.ptr;
			impl->Launch( kkOverflow, bytes, [&]( hipEvent_t e0, hipEvent_t e1 ) {
				if ( L.overflowPlan.empty() == false )
				{
					const dim3 g( 1 ), b( kOverflowThreads );
					const size_t lds = OverflowLdsBytes();
					switch ( stage )
					{
						case kStageWarmStart:
							hipExtLaunchKernelGGL( OverflowLevelKernel<kStageWarmStart>, g, b, lds, s, e0, e1, 0, ca, plan );
							break;
						case kStageSolve:
							hipExtLaunchKernelGGL( OverflowLevelKernel<kStageSolve>, g, b, lds, s, e0, e1, 0, ca, plan );
							break;
						case kStageRelax:
							hipExtLaunchKernelGGL( OverflowLevelKernel<kStageRelax>, g, b, lds, s, e0, e1, 0, ca, plan );
							break;
						default:
							hipExtLaunchKernelGGL( OverflowLevelKernel<kStageRestitution>, g, b, lds, s, e0, e1, 0, ca, plan );
							break;
					}
					return;
				}
				switch ( stage )
				{
					case kStageWarmStart:
						hipExtLaunchKernelGGL( OverflowKernel<kStageWarmStart>, dim3( 1 ), dim3( 64 ), 0, s, e0, e1, 0, ca );
						break;
					case kStageSolve:
						hipExtLaunchKernelGGL( OverflowKernel<kStageSolve>, dim3( 1 ), dim3( 64 ), 0, s, e0, e1, 0, ca );
						break;
					case kStageRelax:
						hipExtLaunchKernelGGL( OverflowKernel<kStageRelax>, dim3( 1 ), dim3( 64 ), 0, s, e0, e1, 0, ca );
						break;
					default:
						hipExtLaunchKernelGGL( OverflowKernel<kStageRestitution>, dim3( 1 ), dim3( 64 ), 0, s, e0, e1, 0, ca );
						break;
				}
			} );
		}

		// the trailing colors of at most kTailThreads lanes each go in one single-workgroup launch
		auto stageLanes = [&]( int c ) {
			int joints = L.colorJointStart[c + 1] - L.colorJointStart[c];
			int contacts = L.colorContactStart[c + 1] - L.colorContactStart[c];
			return stage == kStageRestitution ? ( contacts == 0 ? 0 : joints + contacts ) : joints + contacts;
		};
		int tail = L.activeColorCount;
		while ( tail > 0 && stageLanes( tail - 1 ) <= kTailThreads )
		{
			--tail;
		}
		int launched = 0;
		for ( int c = tail; c < L.activeColorCount; ++c )
		{
			launched += stageLanes( c ) > 0 ? 1 : 0;
		}
		if ( launched < 2 || L.activeColorCount - tail > kTailMaxColors )
		{
			tail = L.activeColorCount;
		}

		for ( int c = 0; c < tail; ++c )
		{
			ca.stage = stage;
			ca.jointStart = L.colorJointStart[c];
			ca.jointCount = L.colorJointStart[c + 1] - L.colorJointStart[c];
			ca.contactStart = L.colorContactStart[c];
			ca.contactCount = L.colorContactStart[c + 1] - L.colorContactStart[c];
			int lanes = ca.jointCount + ca.contactCount;
			if ( lanes == 0 )
			{
				continue;
			}
			if ( stage == kStageRestitution && ca.contactCount == 0 )
			{
				continue;
			}
			double bytes = stageBytes( stage ) * ca.contactCount + kBytesJoint * ca.jointCount;
			impl->Launch( stageClass( stage ), bytes, [&]( hipEvent_t e0, hipEvent_t e1 ) {
				switch ( stage )
				{
					case kStageWarmStart:
						hipExtLaunchKernelGGL( ColorKernel<kStageWarmStart>, dim3( ( lanes + cb - 1 ) / cb ), dim3( cb ), 0, s, e0, e1, 0, ca );
						break;
					case kStageSolve:
						hipExtLaunchKernelGGL( ColorKernel<kStageSolve>, dim3( ( lanes + cb - 1 ) / cb ), dim3( cb ), 0, s, e0, e1, 0, ca );
						break;
					case kStageRelax:
						hipExtLaunchKernelGGL( ColorKernel<kStageRelax>, dim3( ( lanes + cb - 1 ) / cb ), dim3( cb ), 0, s, e0, e1, 0, ca );
						break;
					default:
						hipExtLaunchKernelGGL( ColorKernel<kStageRestitution>, dim3( ( lanes + cb - 1 ) / cb ), dim3( cb ), 0, s, e0, e1, 0, ca );
						break;
				}
			} );
		}

		if ( tail < L.activeColorCount )
		{
			TailArgs ta;
			ta.color = ca;
			ta.color.stage = stage;
			ta.colorCount = L.activeColorCount - tail;
			double bytes = 0.0;
			for ( int k = 0; k <= ta.colorCount; ++k )
			{
				ta.jointStart[k] = L.colorJointStart[tail + k];
				ta.contactStart[k] = L.colorContactStart[tail + k];
			}
			for ( int k = 0; k < ta.colorCount; ++k )
			{
				if ( stageLanes( tail + k ) > 0 )
				{
					bytes += stageBytes( stage ) * ( ta.contactStart[k + 1] - ta.contactStart[k] ) + kBytesJoint * ( ta.jointStart[k + 1] - ta.jointStart[k] );
				}
			}
			impl->Launch( kkColorTail, bytes, [&]( hipEvent_t e0, hipEvent_t e1 ) {
				switch ( stage )
				{
					case kStageWarmStart:
						hipExtLaunchKernelGGL( ColorTailKernel<kStageWarmStart>, dim3( 1 ), dim3( kTailThreads ), 0, s, e0, e1, 0, ta );
						break;
					case kStageSolve:
						hipExtLaunchKernelGGL( ColorTailKernel<kStageSolve>, dim3( 1 ), dim3( kTailThreads ), 0, s, e0, e1, 0, ta );
						break;
					case kStageRelax:
						hipExtLaunchKernelGGL( ColorTailKernel<kStageRelax>, dim3( 1 ), dim3( kTailThreads ), 0, s, e0, e1, 0, ta );
						break;
					default:
						hipExtLaunchKernelGGL( ColorTailKernel<kStageRestitution>, dim3( 1 ), dim3( kTailThreads ), 0, s, e0, e1, 0, ta );
						break;
				}
			} );
		}
	};

	// The fused solver runs on one compute unit: it wins while the per-color launch latency of the
	// color kernels dominates, and loses once a world has enough constraints to need the whole GPU
	// or overflow constraints that its single lane walks from device memory.
	bool fits = nb > 0 && nb <= kFusedBodyCapacity && L.activeColorCount <= kMaxColorSlots;
	bool fused = fits && ( impl->solverPath == 2 || ( impl->solverPath == 0 && AutoFused( nc, L.overflowContactCount + L.overflowJointCount ) ) );
	// per-body warm-start lists exist when the host expected this path (see WantsWarmLists)
	bool haveWarmLists = (int)L.warmStart.size() == nb + 1;

	if ( impl->tape != nullptr && nb > 0 )
	{
		if ( fits == false )
		{
			Fatal( "step recording needs the world's awake bodies in one workgroup's LDS (%d awake, at most %d)", nb,
				   kFusedBodyCapacity );
		}
		fused = true;

		TapeStore::WorldTape& wt = impl->tape->worlds[impl->tapeIndex];
		char* dst = impl->tape->Alloc( TapeRecordBytes( nb, nc ) );
		wt.steps.push_back( dst );
		wt.bodyCounts.push_back( nb );
		wt.constraintCounts.push_back( nc );
		wt.subStepCounts.push_back( params.subStepCount );

		TapeRecordArgs ta;
		memset( &ta.header, 0, sizeof( ta.header ) );
		ta.header.bodyCount = nb;
		ta.header.constraintCount = nc;
		ta.header.jointCount = nj;
		ta.header.activeColorCount = L.activeColorCount;
		for ( int c = 0; c <= L.activeColorCount && c < 25; ++c )
		{
			ta.header.colorStart[c] = L.colorContactStart[c];
		}
		ta.header.overflowStart = L.overflowContactStart;
		ta.header.overflowCount = L.overflowContactCount;
		ta.header.subStepCount = params.subStepCount;
		ta.header.timeStep = params.dt;
		ta.header.h = params.h;
		ta.header.inv_h = params.inv_h;
		ta.header.contactSpeed = params.contactSpeed;
		ta.header.maxLinearSpeed = params.maxLinearVelocity;
		ta.header.maxAngularSpeed = ( 0.25f * B2_PI ) * params.inv_dt;
		ta.header.gravity = params.gravity;
		ta.header.enableWarmStarting = params.enableWarmStarting;
		ta.header.inv_dt = params.inv_dt;
		ta.dst = dst;
		ta.constraintsOffset = TapeConstraintsOffset( nb );
		ta.c = impl->soa;
		ta.contacts = impl->contacts.ptr;
		ta.sims = impl->sims.ptr;
		ta.states = impl->states.ptr;
		ta.awake = impl->awakeBodies.ptr;
		ta.bodyCount = nb;
		ta.constraintCount = nc;
		int lanes = std::max( nb, nc );
		if ( rec )
		{
			impl->rec.tapeRecord = true;
			impl->rec.tapeArgs = ta;
			impl->rec.tapeLanes = lanes;
		}
		else
		{
			LaneKernel<TapeRecordOp><<<Blocks( lanes ), kBlock, 0, s>>>( ta );
		}
	}
	if ( rec && nb > 0 && fits == false )
	{
		Fatal( "batched stepping needs every world's awake bodies in one workgroup's LDS (%d awake, at most %d)", nb,
			   kFusedBodyCapacity );
	}
	if ( rec && nb > 0 )
	{
		fused = true;
	}
	// The XCD solver when asked for, for worlds without overflow constraints. Auto keeps the color
	// kernels: on MI355X both spend about 3.5-4.5 us per color phase (the phases are bound by the
	// gather-solve-scatter latency chain, not by the launch), and the XCD solver measured 0.46 vs
	// 0.42 ms/step on the large pyramid and 1.73 vs 1.82 ms/step on the 20k-body rain.
	bool noOverflow = L.overflowContactCount == 0 && L.overflowJointCount == 0;
	bool xcd = fused == false && rec == false && impl->deferredLaunched == false && nb > 0 && noOverflow &&
			   L.activeColorCount <= kMaxColorSlots && impl->solverPath == 3;
	if ( xcd && impl->xcdUsable < 0 )
	{
		impl->xcdUsable = XcdSolverUsable() ? 1 : 0;
		if ( impl->xcdUsable == 1 )
		{
			HIP_CHECK( hipMalloc( &impl->xcdSync, 4 * sizeof( int ) ) );
		}
	}
	xcd = xcd && impl->xcdUsable == 1;

	if ( xcd )
	{
		XcdArgs xa;
		FusedArgs& fa = xa.f;
		fa.color = ca; // body ids (GlobalBodies)
		fa.awake = impl->awakeBodies.ptr;
		fa.bodyCount = nb;
		fa.sims = impl->sims.ptr;
		fa.states = impl->states.ptr;
		fa.activeColorCount = L.activeColorCount;
		for ( int c = 0; c <= L.activeColorCount; ++c )
		{
			fa.colorContactStart[c] = L.colorContactStart[c];
			fa.colorJointStart[c] = L.colorJointStart[c];
		}
		fa.overflowContactStart = L.overflowContactStart;
		fa.overflowContactCount = 0;
		fa.overflowJointStart = L.overflowJointStart;
		fa.overflowJointCount = 0;
		fa.subStepCount = params.subStepCount;
		fa.anyRestitution = anyRestitution ? 1 : 0;
		// per body when the world has no joints (joint rows would cross workgroups)
		fa.warmStart = impl->warmStart.ptr;
		fa.warmList = impl->warmList.ptr;
		fa.warmByBody = WarmByBody( false ) && haveWarmLists && nj == 0 ? 1 : 0;
		xa.sync = impl->xcdSync;
		xa.zero = 0;
		double bytes = params.subStepCount * ( ( kBytesIntegrateVelocities + kBytesIntegratePositions ) * nb +
											   ( kBytesWarmStart + kBytesSolve + kBytesRelax ) * nc + 3.0 * kBytesJoint * nj ) +
					   ( anyRestitution ? kBytesRestitution * nc : 0.0 );
		HIP_CHECK( hipMemsetAsync( impl->xcdSync, 0, 4 * sizeof( int ), s ) );
		impl->Launch( kkSolveXcd, bytes, [&]( hipEvent_t e0, hipEvent_t e1 ) {
			if ( e0 != nullptr )
			{
				HIP_CHECK( hipEventRecord( e0, s ) );
			}
			void* args[] = { &xa };
			HIP_CHECK( hipLaunchCooperativeKernel( reinterpret_cast<const void*>( SolveXcdKernel ), dim3( kXcdGrid ), dim3( kXcdThreads ), args,
												   0, s ) );
			if ( e1 != nullptr )
			{
				HIP_CHECK( hipEventRecord( e1, s ) );
			}
		} );
	}
	else if ( fused )
	{
		FusedArgs fa;
		fa.color = ca;
		fa.color.c.bodyA = impl->soa.slotA;
		fa.color.c.bodyB = impl->soa.slotB;
		fa.awake = impl->awakeBodies.ptr;
		fa.bodyCount = nb;
		fa.sims = impl->sims.ptr;
		fa.states = impl->states.ptr;
		fa.activeColorCount = L.activeColorCount;
		for ( int c = 0; c <= L.activeColorCount; ++c )
		{
			fa.colorContactStart[c] = L.colorContactStart[c];
			fa.colorJointStart[c] = L.colorJointStart[c];
		}
		fa.overflowContactStart = L.overflowContactStart;
		fa.overflowContactCount = L.overflowContactCount;
		fa.overflowJointStart = L.overflowJointStart;
		fa.overflowJointCount = L.overflowJointCount;
		fa.subStepCount = params.subStepCount;
		fa.anyRestitution = anyRestitution ? 1 : 0;
		fa.warmStart = impl->warmStart.ptr;
		fa.warmList = impl->warmList.ptr;
		fa.warmByBody = ( rec ? BatchWarmByBody( nb ) : WarmByBody( true ) ) && haveWarmLists ? 1 : 0;
		// constraint fields per substep (warm start, solve, relax) + body sims per substep + state in/out
		double bytes = params.subStepCount * ( kBytesFusedContactSubstep * nc + kBytesFusedBodySubstep * nb + kBytesJoint * nj ) +
					   kBytesFusedBodyStep * nb;
		if ( rec )
		{
			impl->rec.fused = true;
			impl->rec.fusedArgs = fa;
		}
		else
		{
			impl->Launch( kkSolveFused, bytes, [&]( hipEvent_t e0, hipEvent_t e1 ) { hipExtLaunchKernelGGL( SolveFusedKernel, dim3( 1 ), dim3( kFusedThreads ), FusedLdsBytes( nb ), s, e0, e1, 0, fa ); } );
		}
	}
	else
	{
		WarmArgs wa;
		wa.b = ba;
		wa.c = ca.c;
		wa.joints = impl->joints.ptr;
		wa.jointList = impl->jointList.ptr;
		wa.warmStart = impl->warmStart.ptr;
		wa.warmList = impl->warmList.ptr;
		wa.overflowContactStart = L.overflowContactStart;
		bool byBody = WarmByBody( false ) && haveWarmLists;
		for ( int sub = 0; sub < params.subStepCount; ++sub )
		{
			if ( byBody )
			{
				impl->Launch( kkIntegrateVelocities, kBytesIntegrateVelocities * nb + kBytesWarmStart * nc + kBytesJoint * nj,
							  [&]( hipEvent_t e0, hipEvent_t e1 ) { hipExtLaunchKernelGGL( IntegrateWarmStartKernel, dim3( ( nb + kWarmBlock - 1 ) / kWarmBlock ), dim3( kWarmBlock ), 0, s, e0, e1, 0, wa ); } );
			}
			else
			{
				impl->Launch( kkIntegrateVelocities, kBytesIntegrateVelocities * nb,
							  [&]( hipEvent_t e0, hipEvent_t e1 ) { hipExtLaunchKernelGGL( IntegrateVelocitiesKernel, dim3( Blocks( nb ) ), dim3( kBlock ), 0, s, e0, e1, 0, ba ); } );
				runStage( kStageWarmStart );
			}
			runStage( kStageSolve );
			impl->Launch( kkIntegratePositions, kBytesIntegratePositions * nb,
						  [&]( hipEvent_t e0, hipEvent_t e1 ) { hipExtLaunchKernelGGL( IntegratePositionsKernel, dim3( ( nb + kWarmBlock - 1 ) / kWarmBlock ), dim3( kWarmBlock ), 0, s, e0, e1, 0, ba ); } );
			runStage( kStageRelax );
		}

		if ( anyRestitution )
		{
			runStage( kStageRestitution );
		}
	}

	// Without recording, store, finalize, the continuous sweeps and the pack run as one launch
	// (FinishKernel); B2AMD_NO_FUSED_FINISH keeps them separate launches.
	static const bool fuseFinishAllowed = getenv( "B2AMD_NO_FUSED_FINISH" ) == nullptr;
	const bool fuseFinish = rec == false && fuseFinishAllowed;
	FinishArgs fin = {};
	StoreArgs sa = {};
	if ( nc > 0 )
	{
		sa.c = impl->soa;
		sa.contacts = impl->contacts.ptr;
		sa.coloredCount = L.overflowContactStart;
		sa.totalCount = nc;
		impl->hitCandidates.Reserve( std::max<size_t>( nc, 1 ), false, s, 0 );
		sa.hitCandidates = impl->hitCandidates.ptr;
		sa.hitCount = impl->counters + 8;
		sa.negHitThreshold = -params.hitEventThreshold;
		if ( rec )
		{
			impl->rec.store = true;
			impl->rec.storeArgs = sa;
			impl->rec.storeLanes = nc;
		}
		else if ( fuseFinish == false )
		{
			impl->Launch( kkStore, kBytesStore * nc, [&]( hipEvent_t e0, hipEvent_t e1 ) { hipExtLaunchKernelGGL( LaneKernel<StoreImpulsesOp>, dim3( Blocks( nc ) ), dim3( kBlock ), 0, s, e0, e1, 0, sa ); } );
		}
	}

	FinalizeArgs fa = {};
	fa.awake = impl->awakeBodies.ptr;
	fa.count = nb;
	fa.sims = impl->sims.ptr;
	fa.states = impl->states.ptr;
	fa.aux = impl->aux.ptr;
	fa.shapes = impl->shapes.ptr;
	fa.p = params;
	fa.counters = impl->counters + 4;
	fa.fastBodies = impl->fastBodies.ptr;
	fa.enlargedBodies = impl->enlargedBodies.ptr;
	fa.enlargedShapes = impl->enlargedShapes.ptr;
	fa.enlargedShapeAABBs = impl->enlargedShapeAABBs.ptr;
	const bool islands = needSleepTimes && impl->islandCount > 0;
	fa.islandSlots = islands ? impl->islandSlots.ptr : nullptr;
	fa.islandIds = islands ? impl->islandIds.ptr : nullptr;
	fa.islandBits = islands ? impl->islandBits.ptr : nullptr;
	fa.splitKey = reinterpret_cast<unsigned long long*>( impl->counters + kSplitKeySlot );
	if ( rec )
	{
		impl->rec.finalize = true;
		impl->rec.finalizeArgs = fa;
		impl->rec.finalizeLanes = nb;
	}
	else if ( fuseFinish == false )
	{
		impl->Launch( kkFinalize, kBytesFinalize * nb, [&]( hipEvent_t e0, hipEvent_t e1 ) { hipExtLaunchKernelGGL( LaneKernel<FinalizeBodiesOp>, dim3( Blocks( nb ) ), dim3( kBlock ), 0, s, e0, e1, 0, fa ); } );
	}

	if ( params.enableContinuous && params.hostContinuous == 0 )
	{
		// The fast-body count is only known on the device: size the grid for every awake body.
		ContinuousArgs ka = {};
		ka.fastBodies = impl->fastBodies.ptr;
		ka.fastCount = impl->counters + 6;
		ka.awake = impl->awakeBodies.ptr;
		ka.sims = impl->sims.ptr;
		ka.shapes = impl->shapes.ptr;
		ka.tree = impl->ccdTree.ptr;
		ka.treeRoot = impl->ccdTree.ptr != nullptr ? impl->ccdTreeRoot : B2_NULL_INDEX;
		ka.error = impl->counters + 7;
		ka.noCollide = impl->ccdNoCollide.ptr;
		ka.noCollideCount = (int)impl->ccdNoCollideStaging.size();
		ka.filterRejects = impl->ccdFilterRejects.ptr;
		ka.filterRejectCount = impl->ccdFilterRejectCount;
		ka.enlargedShapeCount = impl->counters + 5;
		ka.enlargedShapes = impl->enlargedShapes.ptr;
		ka.enlargedShapeAABBs = impl->enlargedShapeAABBs.ptr;
		ka.huge = B2_HUGE;
		// at most kMaxContinuousSensorHits per fast body
		impl->sensorHitCapacity = params.sensorHits != 0 ? kMaxContinuousSensorHits * std::max( nb, 1 ) : 0;
		if ( impl->sensorHitCapacity > 0 )
		{
			impl->sensorHits.Reserve( impl->sensorHitCapacity, false, s, 0 );
		}
		ka.sensorHits = impl->sensorHitCapacity > 0 ? impl->sensorHits.ptr : nullptr;
		ka.sensorHitCount = impl->counters + kSensorHitCountSlot;
		ka.sensorHitCapacity = impl->sensorHitCapacity;
		if ( rec )
		{
			impl->rec.continuous = true;
			impl->rec.continuousArgs = ka;
			impl->rec.continuousLanes = nb;
		}
		else if ( fuseFinish )
		{
			fin.ccd = ka;
			fin.sweep = 1;
		}
		else
		{
			impl->Launch( kkContinuous, 0.0, [&]( hipEvent_t e0, hipEvent_t e1 ) { hipExtLaunchKernelGGL( LaneKernel<ContinuousOp>, dim3( Blocks( nb ) ), dim3( kBlock ), 0, s, e0, e1, 0, ka ); } );
		}
	}
	else
	{
		impl->sensorHitCapacity = 0;
	}
	if ( rec == false && fuseFinish == false )
	{
		HIP_CHECK( hipEventRecord( impl->ev[3], s ) );
	}

	{
		size_t shapeCap = std::max<size_t>( impl->shapeCount, 1 );
		PackFinalizeArgs pk = {};
		pk.counters = impl->counters + 4;
		pk.enlargedBodies = impl->enlargedBodies.ptr;
		pk.enlargedShapes = impl->enlargedShapes.ptr;
		pk.enlargedShapeAABBs = impl->enlargedShapeAABBs.ptr;
		pk.islandBits = islands ? impl->islandBits.ptr : nullptr;
		pk.splitKey = reinterpret_cast<unsigned long long*>( impl->counters + kSplitKeySlot );
		pk.islandWords = islands ? ( impl->islandCount + 31 ) / 32 : 0;
		pk.keptShapeCount = rec ? nullptr : impl->counters + kKeptShapeCountSlot;
		pk.xcdSync = xcd ? impl->xcdSync : nullptr;
		pk.bodyCount = nb;
		pk.shapeOffset = kMailHeader + nb;
		pk.aabbOffset = pk.shapeOffset + (int)shapeCap;
		pk.sleepOffset = pk.aabbOffset + 4 * (int)shapeCap;
		pk.hitOffset = pk.sleepOffset + std::max( nb, pk.islandWords + 2 );
		pk.joints = impl->joints.ptr;
		pk.jointCount = params.jointEvents != 0 ? impl->jointCount : 0;
		pk.jointEventOffset = pk.hitOffset + kHitRecordInts * std::max( nc, 0 );
		impl->solveHitCapacity = nc;
		impl->hitCandidates.Reserve( std::max<size_t>( nc, 1 ), false, s, 0 );
		pk.contacts = impl->contacts.ptr;
		pk.sims = impl->sims.ptr;
		pk.aux = impl->aux.ptr;
		pk.hitCandidates = impl->hitCandidates.ptr;
		pk.hitCount = impl->counters + 8;
		pk.hitThreshold = params.hitEventThreshold;
		impl->EnsureMail( (size_t)pk.jointEventOffset + (size_t)pk.jointCount );
		pk.mail = impl->mail;
		impl->solveBodyCount = nb;
		impl->solveNeedSleepTimes = needSleepTimes;
		impl->solvePack = pk;
		if ( rec )
		{
			impl->rec.packFinalize = true;
			impl->rec.packFinalizeArgs = pk;
		}
		else if ( fuseFinish )
		{
			fin.store = sa;
			fin.fin = fa;
			fin.pack = pk;
			// the lanes write the enlarged lists straight into the mailbox
			fin.fin.mailBodies = impl->mail + kMailHeader;
			fin.fin.mailShapes = impl->mail + pk.shapeOffset;
			fin.fin.mailAabbs = reinterpret_cast<float*>( impl->mail + pk.aabbOffset );
			fin.ccd.mailShapes = fin.fin.mailShapes;
			fin.ccd.mailAabbs = fin.fin.mailAabbs;
			fin.pack.listsInMail = 1;
			fin.storeBlocks = ( nc + kFinishThreads - 1 ) / kFinishThreads;
			fin.done = impl->counters + kFinishDoneSlot;
			const int blocks = std::max( 1, fin.storeBlocks + ( nb + kFinishThreads - 1 ) / kFinishThreads );
			impl->Launch( kkFinish, kBytesStore * nc + kBytesFinalize * nb, [&]( hipEvent_t e0, hipEvent_t e1 ) {
				hipExtLaunchKernelGGL( FinishKernel, dim3( blocks ), dim3( kFinishThreads ), 0, s, e0, e1, 0, fin );
			} );
			HIP_CHECK( hipEventRecord( impl->ev[3], s ) );
			HIP_CHECK( hipEventRecord( impl->solvePacked, s ) );
			impl->solvePackedValid = true;
		}
		else
		{
			PackFinalizeKernel<<<1, kPackThreads, 0, s>>>( pk );
			HIP_CHECK( hipGetLastError() );
			HIP_CHECK( hipEventRecord( impl->solvePacked, s ) );
			impl->solvePackedValid = true;
		}
	}
}

bool DeviceWorld::GetSolvePairInputs( SolvePairInputs* out ) const
{
	JoinLaunch();
	if ( impl->recording || impl->solvePackedValid == false )
	{
		return false;
	}
	out->enlargedShapes = impl->enlargedShapes.ptr;
	out->enlargedShapeAABBs = impl->enlargedShapeAABBs.ptr;
	out->enlargedShapeCount = impl->counters + kKeptShapeCountSlot;
	out->capacity = (int)impl->enlargedShapes.capacity;
	out->ready = impl->solvePacked;
	return true;
}

void DeviceWorld::UploadIslands( const std::vector<int>& slots, const std::vector<int>& ids, int islandCount )
{
	impl->StageVector( impl->islandSlots, slots );
	impl->StageVector( impl->islandIds, ids );
	size_t words = ( (size_t)islandCount + 31 ) / 32;
	impl->islandBits.Reserve( std::max<size_t>( words, 1 ), false, impl->stream, 0 );
	if ( impl->islandBits.capacity > impl->islandBitsZeroed )
	{
		// the pack kernel clears the words it reads; fresh storage starts at zero once
		HIP_CHECK( hipMemsetAsync( impl->islandBits.ptr, 0, impl->islandBits.capacity * sizeof( uint32_t ), impl->stream ) );
		impl->islandBitsZeroed = impl->islandBits.capacity;
	}
	impl->islandCount = islandCount;
}

void DeviceWorld::LaunchDeferredSolve()
{
	if ( impl->solveDeferred == false )
	{
		return;
	}
	impl->solveDeferred = false;
	impl->deferredLaunched = true;
	impl->recording = false;
	SolveLaunch( impl->deferredParams, impl->deferredRestitution, impl->deferredSleepTimes );
	impl->recording = true;
}

void DeviceWorld::SolveFinish( FinalizeResult* result )
{
	JoinLaunch();
	const bool ownStream = impl->recording == false || impl->deferredLaunched;
	impl->deferredLaunched = false;
	{
		int nb = impl->solveBodyCount;
		bool needSleepTimes = impl->solveNeedSleepTimes;
		const PackFinalizeArgs& pk = impl->solvePack;
		if ( ownStream )
		{
			HIP_CHECK( hipStreamSynchronize( impl->stream ) );
		}
		impl->FlushTiming();

		const int* mail = impl->mail;
		int enlargedBodyCount = mail[0];
		int enlargedShapeCount = mail[1];
		result->fastBodyCount = mail[2];
		if ( mail[3] != 0 )
		{
			Fatal( "continuous collision: the static tree is deeper than the query stack (%d entries)", kCcdStackSize );
		}
		if ( mail[5] != 0 )
		{
			Fatal( "XCD solver: a barrier among its workgroups timed out (set B2AMD_NO_XCD_SOLVER to use the color kernels)" );
		}
		result->enlargedBodies.assign( mail + kMailHeader, mail + kMailHeader + enlargedBodyCount );
		result->enlargedShapes.assign( mail + pk.shapeOffset, mail + pk.shapeOffset + enlargedShapeCount );
		result->enlargedShapeAABBs.resize( enlargedShapeCount );
		if ( enlargedShapeCount > 0 )
		{
			memcpy( result->enlargedShapeAABBs.data(), mail + pk.aabbOffset, enlargedShapeCount * sizeof( b2AABB ) );
		}
		result->islandBits.clear();
		result->splitKey = 0;
		if ( pk.islandBits != nullptr )
		{
			const uint32_t* bits = reinterpret_cast<const uint32_t*>( mail + pk.sleepOffset );
			result->islandBits.assign( bits, bits + pk.islandWords );
			result->splitKey = (uint64_t)bits[pk.islandWords] | (uint64_t)bits[pk.islandWords + 1] << 32;
		}
		std::sort( result->enlargedBodies.begin(), result->enlargedBodies.end() );

		int hitCount = mail[4];
		if ( hitCount > impl->solveHitCapacity )
		{
			Fatal( "hit event count %d exceeds the colored contact count %d", hitCount, impl->solveHitCapacity );
		}
		result->hits.clear();
		for ( int h = 0; h < hitCount; ++h )
		{
			const int* rec = mail + pk.hitOffset + kHitRecordInts * h;
			if ( rec[0] < 0 )
			{
				continue;
			}
			HitRecord hr;
			hr.contactId = rec[0];
			hr.shapeIdA = rec[1];
			hr.shapeIdB = rec[2];
			memcpy( &hr.point, rec + 3, 2 * sizeof( float ) );
			memcpy( &hr.normal, rec + 5, 2 * sizeof( float ) );
			memcpy( &hr.approachSpeed, rec + 7, sizeof( float ) );
			result->hits.push_back( hr );
		}
		// the reference reports hits in ascending contact id
		std::sort( result->hits.begin(), result->hits.end(), []( const HitRecord& x, const HitRecord& y ) { return x.contactId < y.contactId; } );

		// and joint events in ascending joint id
		int jointEventCount = mail[6];
		if ( jointEventCount < 0 || jointEventCount > pk.jointCount )
		{
			Fatal( "joint event count %d exceeds the joint count %d", jointEventCount, pk.jointCount );
		}
		result->jointEvents.assign( mail + pk.jointEventOffset, mail + pk.jointEventOffset + jointEventCount );
		std::sort( result->jointEvents.begin(), result->jointEvents.end() );
	}

	if ( ownStream )
	{
		HIP_CHECK( hipEventElapsedTime( &lastSolveMs, impl->ev[2], impl->ev[3] ) );
	}
}

void DeviceWorld::Solve( const StepParams& params, bool anyRestitution, bool needSleepTimes, FinalizeResult* result )
{
	SolveLaunch( params, anyRestitution, needSleepTimes );
	SolveFinish( result );
}

void DeviceWorld::EnableKernelTiming( bool flag )
{
	impl->timing = flag;
}

int DeviceWorld::GetKernelStats( void* out, int capacity )
{
	b2amd_KernelStats* stats = static_cast<b2amd_KernelStats*>( out );
	int n = 0;
	for ( int k = 0; k < kKernelClassCount && n < capacity; ++k )
	{
		if ( impl->statLaunches[k] == 0 )
		{
			continue;
		}
		b2amd_KernelStats& st = stats[n++];
		memset( &st, 0, sizeof( st ) );
		strncpy( st.name, s_kernelClassNames[k], sizeof( st.name ) - 1 );
		st.launches = impl->statLaunches[k];
		st.totalMs = impl->statMs[k];
		st.algorithmicBytes = impl->statBytes[k];
	}
	return n;
}

void DeviceWorld::ResetKernelStats()
{
	for ( int k = 0; k < kKernelClassCount; ++k )
	{
		impl->statMs[k] = 0.0;
		impl->statBytes[k] = 0.0;
		impl->statLaunches[k] = 0;
	}
}

// ---------------------------------------------------------------------------------------------
// Batched narrowphase for callers outside the world step.

__global__ void CollideBatchKernel( const ShapeDev* shapesA, const ShapeDev* shapesB, const b2Transform* xfs, int count,
									b2LocalManifold* out )
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if ( i >= count )
	{
		return;
	}
	out[i] = ComputeManifold( shapesA[i], shapesB[i], xfs[i] );
}

void DeviceWorld::SetSolverPath( int path )
{
	impl->solverPath = path;
}

bool DeviceWorld::WantsWarmLists( int bodyCount, int constraintCount, int activeColorCount, int overflowCount ) const
{
	if ( bodyCount == 0 )
	{
		return false;
	}
	bool fits = bodyCount <= kFusedBodyCapacity && activeColorCount <= kMaxColorSlots;
	bool fused = fits && ( impl->solverPath == 2 || ( impl->solverPath == 0 && AutoFused( constraintCount, overflowCount ) ) );
	if ( impl->recording )
	{
		return BatchWarmByBody( bodyCount );
	}
	fused = fused || impl->tape != nullptr;
	return WarmByBody( fused );
}

void DeviceWorld::SetTape( TapeStore* tape, int index )
{
	impl->tape = tape;
	impl->tapeIndex = index;
}

TapeStore* DeviceWorld::Tape() const
{
	return impl->tape;
}

void DeviceWorld::CaptureFinalState( char* dst, int bodyCount )
{
	TapeRecordArgs ta;
	memset( &ta, 0, sizeof( ta ) );
	ta.header.bodyCount = bodyCount;
	ta.dst = dst;
	ta.constraintsOffset = TapeConstraintsOffset( bodyCount );
	ta.c = impl->soa;
	ta.contacts = impl->contacts.ptr;
	ta.sims = impl->sims.ptr;
	ta.states = impl->states.ptr;
	ta.awake = impl->awakeBodies.ptr;
	ta.bodyCount = bodyCount;
	ta.constraintCount = 0;
	int lanes = std::max( bodyCount, 1 );
	LaneKernel<TapeRecordOp><<<Blocks( lanes ), kBlock, 0, impl->stream>>>( ta );
	HIP_CHECK( hipGetLastError() );
	HIP_CHECK( hipStreamSynchronize( impl->stream ) );
}

const void* DeviceWorld::ShapesDevice() const
{
	return impl->shapes.ptr;
}

const void* DeviceWorld::SimsDevice() const
{
	return impl->sims.ptr;
}

int DeviceWorld::AwakeBodyCount() const
{
	return (int)impl->lists.awakeBodies.size();
}

__global__ void TimeOfImpactBatchKernel( const b2TOIInput* inputs, int count, b2TOIOutput* outputs )
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if ( i < count )
	{
		outputs[i] = TimeOfImpact( inputs[i] );
	}
}

__global__ void ShapeDistanceBatchKernel( const b2DistanceInput* inputs, b2SimplexCache* caches, int count, b2DistanceOutput* outputs )
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if ( i < count )
	{
		b2SimplexCache cache = caches[i];
		outputs[i] = ShapeDistance( inputs[i], &cache );
		caches[i] = cache;
	}
}

// Copy n records to a fresh device buffer.
template <typename T>
static T* ToDevice( const T* host, int n )
{
	T* d = nullptr;
	HIP_CHECK( hipMalloc( &d, std::max( n, 1 ) * sizeof( T ) ) );
	HIP_CHECK( hipMemcpy( d, host, n * sizeof( T ), hipMemcpyHostToDevice ) );
	return d;
}

int TimeOfImpactBatch( const b2TOIInput* inputs, int count, b2TOIOutput* outputs )
{
	if ( count <= 0 )
	{
		return 0;
	}
	b2TOIInput* dIn = ToDevice( inputs, count );
	b2TOIOutput* dOut = nullptr;
	HIP_CHECK( hipMalloc( &dOut, count * sizeof( b2TOIOutput ) ) );
	TimeOfImpactBatchKernel<<<Blocks( count ), kBlock>>>( dIn, count, dOut );
	HIP_CHECK( hipGetLastError() );
	HIP_CHECK( hipMemcpy( outputs, dOut, count * sizeof( b2TOIOutput ), hipMemcpyDeviceToHost ) );
	HIP_CHECK( hipFree( dIn ) );
	HIP_CHECK( hipFree( dOut ) );
	return 0;
}

int ShapeDistanceBatch( const b2DistanceInput* inputs, b2SimplexCache* caches, int count, b2DistanceOutput* outputs )
{
	if ( count <= 0 )
	{
		return 0;
	}
	b2DistanceInput* dIn = ToDevice( inputs, count );
	b2SimplexCache* dCache = ToDevice( caches, count );
	b2DistanceOutput* dOut = nullptr;
	HIP_CHECK( hipMalloc( &dOut, count * sizeof( b2DistanceOutput ) ) );
	ShapeDistanceBatchKernel<<<Blocks( count ), kBlock>>>( dIn, dCache, count, dOut );
	HIP_CHECK( hipGetLastError() );
	HIP_CHECK( hipMemcpy( outputs, dOut, count * sizeof( b2DistanceOutput ), hipMemcpyDeviceToHost ) );
	HIP_CHECK( hipMemcpy( caches, dCache, count * sizeof( b2SimplexCache ), hipMemcpyDeviceToHost ) );
	HIP_CHECK( hipFree( dIn ) );
	HIP_CHECK( hipFree( dCache ) );
	HIP_CHECK( hipFree( dOut ) );
	return 0;
}

void DeviceSynchronize()
{
	HIP_CHECK( hipDeviceSynchronize() );
}

bool SpinWaitActive()
{
	unsigned int flags = 0;
	if ( hipGetDeviceFlags( &flags ) != hipSuccess )
	{
		(void)hipGetLastError();
		return false;
	}
	return ( flags & hipDeviceScheduleMask ) == hipDeviceScheduleSpin;
}

// A pair the batched narrowphase accepts: known shape types in the reference's primary order.
static bool CollidePairSupported( int tA, int tB )
{
	bool known = ( tA == b2_circleShape || tA == b2_capsuleShape || tA == b2_segmentShape || tA == b2_polygonShape ) &&
				 ( tB == b2_circleShape || tB == b2_capsuleShape || tB == b2_segmentShape || tB == b2_polygonShape );
	bool primary = ( tA == tB ) || ( tA == b2_capsuleShape && tB == b2_circleShape ) ||
				   ( tA == b2_polygonShape && ( tB == b2_circleShape || tB == b2_capsuleShape ) ) ||
				   ( tA == b2_segmentShape && tB != b2_segmentShape );
	return known && primary && ( tA == b2_segmentShape && tB == b2_segmentShape ) == false;
}

// ---------------------------------------------------------------------------------------------
// Batched narrowphase without allocation (b2amd_CollideBatchNoAlloc).
//
// The pairs and their manifolds are staged in a file-scope device array instead of buffers
// allocated per call. Each GPU holds its own copy of the array for the life of the process; its
// address on a device is looked up once, with that device current, and kept per device. Every
// call on one GPU reuses that GPU's array, so calls that overlap on one GPU race.

constexpr int kNoAllocCollidePairs = 4096;
constexpr int kNoAllocMaxDevices = 64;

struct CollideScratch
{
	b2amd_Shape shapesA[kNoAllocCollidePairs];
	b2amd_Shape shapesB[kNoAllocCollidePairs];
	b2Transform transforms[kNoAllocCollidePairs];
	b2LocalManifold manifolds[kNoAllocCollidePairs];
};

static __device__ __align__( 256 ) CollideScratch scratch;

static CollideScratch* s_scratchByDevice[kNoAllocMaxDevices] = {};
static std::mutex s_scratchMutex;

// The current device's address of the scratch array (looked up on the device's first call).
static CollideScratch* ScratchOfCurrentDevice()
{
	int device = 0;
	if ( hipGetDevice( &device ) != hipSuccess || device < 0 || device >= kNoAllocMaxDevices )
	{
		(void)hipGetLastError();
		Log( "b2amd_CollideBatchNoAlloc: no usable current GPU (device %d)", device );
		return nullptr;
	}
	std::lock_guard<std::mutex> lock( s_scratchMutex );
	if ( s_scratchByDevice[device] == nullptr )
	{
		void* ptr = nullptr;
		hipError_t e = hipGetSymbolAddress( &ptr, HIP_SYMBOL( scratch ) );
		if ( e != hipSuccess || ptr == nullptr )
		{
			(void)hipGetLastError();
			Log( "b2amd_CollideBatchNoAlloc: hipGetSymbolAddress failed on device %d: %s", device, hipGetErrorString( e ) );
			return nullptr;
		}
		s_scratchByDevice[device] = static_cast<CollideScratch*>( ptr );
	}
	return s_scratchByDevice[device];
}

// One lane per pair: the geometry is read in its public form and placed in the world step's
// shape record (all other fields zero, as b2amd_CollideBatch stages it) before the manifold.
__global__ void CollidePublicShapesKernel( const b2amd_Shape* shapesA, const b2amd_Shape* shapesB, const b2Transform* xfs,
										   int count, b2LocalManifold* out )
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if ( i >= count )
	{
		return;
	}
	ShapeDev s[2];
	const b2amd_Shape* src[2] = { shapesA + i, shapesB + i };
	for ( int k = 0; k < 2; ++k )
	{
		memset( &s[k], 0, sizeof( ShapeDev ) );
		s[k].type = src[k]->type;
		switch ( src[k]->type )
		{
			case b2_circleShape:
				s[k].circle = src[k]->circle;
				break;
			case b2_capsuleShape:
				s[k].capsule = src[k]->capsule;
				break;
			case b2_segmentShape:
				s[k].segment = src[k]->segment;
				break;
			default:
				s[k].polygon = src[k]->polygon;
				break;
		}
	}
	out[i] = ComputeManifold( s[0], s[1], xfs[i] );
}

int CollideBatchNoAlloc( const b2amd_Shape* shapesA, const b2amd_Shape* shapesB, const b2Transform* transforms, int count,
						 b2LocalManifold* manifolds )
{
	if ( count <= 0 )
	{
		return 0;
	}
	if ( count > kNoAllocCollidePairs )
	{
		Log( "b2amd_CollideBatchNoAlloc: %d pairs exceed the scratch capacity of %d pairs", count, kNoAllocCollidePairs );
		return -2;
	}
	for ( int i = 0; i < count; ++i )
	{
		if ( CollidePairSupported( shapesA[i].type, shapesB[i].type ) == false )
		{
			return -1;
		}
	}
	CollideScratch* ptr = ScratchOfCurrentDevice();
	if ( ptr == nullptr )
	{
		return -3;
	}
	HIP_CHECK( hipMemcpy( ptr->shapesA, shapesA, count * sizeof( b2amd_Shape ), hipMemcpyHostToDevice ) );
	HIP_CHECK( hipMemcpy( ptr->shapesB, shapesB, count * sizeof( b2amd_Shape ), hipMemcpyHostToDevice ) );
	HIP_CHECK( hipMemcpy( ptr->transforms, transforms, count * sizeof( b2Transform ), hipMemcpyHostToDevice ) );
	CollidePublicShapesKernel<<<Blocks( count ), kBlock>>>( ptr->shapesA, ptr->shapesB, ptr->transforms, count, ptr->manifolds );
	HIP_CHECK( hipGetLastError() );
	HIP_CHECK( hipMemcpy( manifolds, ptr->manifolds, count * sizeof( b2LocalManifold ), hipMemcpyDeviceToHost ) );
	return 0;
}

int CollideBatch( const b2amd_Shape* shapesA, const b2amd_Shape* shapesB, const b2Transform* transforms, int count,
				  b2LocalManifold* manifolds )
{
	if ( count <= 0 )
	{
		return 0;
	}

	std::vector<ShapeDev> a( count ), b( count );
	for ( int i = 0; i < count; ++i )
	{
		const b2amd_Shape* src[2] = { shapesA + i, shapesB + i };
		ShapeDev* dst[2] = { &a[i], &b[i] };
		for ( int k = 0; k < 2; ++k )
		{
			memset( dst[k], 0, sizeof( ShapeDev ) );
			dst[k]->type = src[k]->type;
			switch ( src[k]->type )
			{
				case b2_circleShape:
					dst[k]->circle = src[k]->circle;
					break;
				case b2_capsuleShape:
					dst[k]->capsule = src[k]->capsule;
					break;
				case b2_segmentShape:
					dst[k]->segment = src[k]->segment;
					break;
				case b2_polygonShape:
					dst[k]->polygon = src[k]->polygon;
					break;
				default:
					return -1;
			}
		}
		int tA = a[i].type, tB = b[i].type;
		bool primary = ( tA == tB ) || ( tA == b2_capsuleShape && tB == b2_circleShape ) ||
					   ( tA == b2_polygonShape && ( tB == b2_circleShape || tB == b2_capsuleShape ) ) ||
					   ( tA == b2_segmentShape && tB != b2_segmentShape );
		if ( primary == false || ( tA == b2_segmentShape && tB == b2_segmentShape ) )
		{
			return -1;
		}
	}

	ShapeDev* dA = nullptr;
	ShapeDev* dB = nullptr;
	b2Transform* dX = nullptr;
	b2LocalManifold* dM = nullptr;
	HIP_CHECK( hipMalloc( &dA, count * sizeof( ShapeDev ) ) );
	HIP_CHECK( hipMalloc( &dB, count * sizeof( ShapeDev ) ) );
	HIP_CHECK( hipMalloc( &dX, count * sizeof( b2Transform ) ) );
	HIP_CHECK( hipMalloc( &dM, count * sizeof( b2LocalManifold ) ) );
	HIP_CHECK( hipMemcpy( dA, a.data(), count * sizeof( ShapeDev ), hipMemcpyHostToDevice ) );
	HIP_CHECK( hipMemcpy( dB, b.data(), count * sizeof( ShapeDev ), hipMemcpyHostToDevice ) );
	HIP_CHECK( hipMemcpy( dX, transforms, count * sizeof( b2Transform ), hipMemcpyHostToDevice ) );
	CollideBatchKernel<<<Blocks( count ), kBlock>>>( dA, dB, dX, count, dM );
	HIP_CHECK( hipGetLastError() );
	HIP_CHECK( hipMemcpy( manifolds, dM, count * sizeof( b2LocalManifold ), hipMemcpyDeviceToHost ) );
	HIP_CHECK( hipFree( dA ) );
	HIP_CHECK( hipFree( dB ) );
	HIP_CHECK( hipFree( dX ) );
	HIP_CHECK( hipFree( dM ) );
	return 0;
}


} // namespace b2g

// Host forms of the manifold functions (the same source the collide kernel runs per lane).
extern "C" {

B2_API b2SegmentDistanceResult b2SegmentDistance( b2Vec2 p1, b2Vec2 q1, b2Vec2 p2, b2Vec2 q2 )
{
	return b2g::SegmentDistance( p1, q1, p2, q2 );
}

B2_API b2LocalManifold b2CollideCircles( const b2Circle* circleA, const b2Circle* circleB, b2Transform xf )
{
	return b2g::CollideCircles( *circleA, *circleB, xf );
}

B2_API b2LocalManifold b2CollideCapsuleAndCircle( const b2Capsule* capsuleA, const b2Circle* circleB, b2Transform xf )
{
	return b2g::CollideCapsuleAndCircle( *capsuleA, *circleB, xf );
}

B2_API b2LocalManifold b2CollideSegmentAndCircle( const b2Segment* segmentA, const b2Circle* circleB, b2Transform xf )
{
	return b2g::CollideSegmentAndCircle( *segmentA, *circleB, xf );
}

B2_API b2LocalManifold b2CollidePolygonAndCircle( const b2Polygon* polygonA, const b2Circle* circleB, b2Transform xf )
{
	return b2g::CollidePolygonAndCircle( *polygonA, *circleB, xf );
}

B2_API b2LocalManifold b2CollideCapsules( const b2Capsule* capsuleA, const b2Capsule* capsuleB, b2Transform xf )
{
	return b2g::CollideCapsules( *capsuleA, *capsuleB, xf );
}

B2_API b2LocalManifold b2CollideSegmentAndCapsule( const b2Segment* segmentA, const b2Capsule* capsuleB, b2Transform xf )
{
	return b2g::CollideSegmentAndCapsule( *segmentA, *capsuleB, xf );
}

B2_API b2LocalManifold b2CollidePolygonAndCapsule( const b2Polygon* polygonA, const b2Capsule* capsuleB, b2Transform xf )
{
	return b2g::CollidePolygonAndCapsule( *polygonA, *capsuleB, xf );
}

B2_API b2LocalManifold b2CollidePolygons( const b2Polygon* polygonA, const b2Polygon* polygonB, b2Transform xf )
{
	return b2g::CollidePolygons( *polygonA, *polygonB, xf );
}

B2_API b2LocalManifold b2CollideSegmentAndPolygon( const b2Segment* segmentA, const b2Polygon* polygonB, b2Transform xf )
{
	return b2g::CollideSegmentAndPolygon( *segmentA, *polygonB, xf );
}

} // extern "C"

// Host forms of the distance and time-of-impact routines (the code the continuous kernel runs).
extern "C" {

B2_API b2ShapeProxy b2MakeProxy( const b2Vec2* points, int count, float radius )
{
	return b2g::MakeProxy( points, count, radius );
}

B2_API b2ShapeProxy b2MakeOffsetProxy( const b2Vec2* points, int count, float radius, b2Vec2 position, b2Rot rotation )
{
	count = count < B2_MAX_POLYGON_VERTICES ? count : B2_MAX_POLYGON_VERTICES;
	b2Transform transform = { position, rotation };
	b2ShapeProxy proxy;
	for ( int i = 0; i < count; ++i )
	{
		proxy.points[i] = b2TransformPoint( transform, points[i] );
	}
	proxy.count = count;
	proxy.radius = radius;
	return proxy;
}

B2_API b2DistanceOutput b2ShapeDistance( const b2DistanceInput* input, b2SimplexCache* cache, b2Simplex* simplexes,
										 int simplexCapacity )
{
	int recorded = 0;
	if ( simplexes != nullptr && simplexCapacity > 0 )
	{
		// the reference records the starting simplex (src/distance.c:444-449)
		b2ShapeProxy localB;
		localB.count = input->proxyB.count;
		localB.radius = input->proxyB.radius;
		for ( int i = 0; i < localB.count; ++i )
		{
			localB.points[i] = b2TransformPoint( input->transform, input->proxyB.points[i] );
		}
		simplexes[0] = b2g::SimplexFromCache( *cache, input->proxyA, localB );
		recorded = 1;
	}
	b2DistanceOutput output = b2g::ShapeDistance( *input, cache );
	output.simplexCount = recorded;
	return output;
}

B2_API b2Transform b2GetSweepTransform( const b2Sweep* sweep, float time )
{
	return b2g::SweepTransform( *sweep, time );
}

B2_API b2TOIOutput b2TimeOfImpact( const b2TOIInput* input )
{
	return b2g::TimeOfImpact( *input );
}

} // extern "C"

// Host forms of the chain-segment manifold functions.
extern "C" {

B2_API b2LocalManifold b2CollideChainSegmentAndCircle( const b2ChainSegment* segmentA, const b2Circle* circleB, b2Transform xf )
{
	return b2g::CollideChainSegmentAndCircle( *segmentA, *circleB, xf );
}

B2_API b2LocalManifold b2CollideChainSegmentAndCapsule( const b2ChainSegment* segmentA, const b2Capsule* capsuleB, b2Transform xf,
														b2SimplexCache* cache )
{
	return b2g::CollideChainSegmentAndCapsule( *segmentA, *capsuleB, xf, cache );
}

B2_API b2LocalManifold b2CollideChainSegmentAndPolygon( const b2ChainSegment* segmentA, const b2Polygon* polygonB, b2Transform xf,
														b2SimplexCache* cache )
{
	return b2g::CollideChainSegmentAndPolygon( *segmentA, *polygonB, xf, cache );
}

} // extern "C"

// ---------------------------------------------------------------------------------------------
// Batched stepping of many worlds: one launch per stage over every world's recorded arguments.

namespace b2g
{

void DeviceWorld::SetRecording( bool flag )
{
	// Uploads and contact inits recorded but not taken by a batch run here, in order.
	Impl::Record& r = impl->rec;
	if ( r.uploads.empty() == false || r.init )
	{
		for ( const auto& u : r.uploads )
		{
			HIP_CHECK( hipMemcpyAsync( u.dst, r.payload.data() + u.offset, u.bytes, hipMemcpyHostToDevice, impl->stream ) );
		}
		if ( r.init )
		{
			LaneKernel<InitContactsOp><<<Blocks( r.initLanes ), kBlock, 0, impl->stream>>>( r.initArgs );
			HIP_CHECK( hipGetLastError() );
		}
		HIP_CHECK( hipStreamSynchronize( impl->stream ) );
	}
	impl->recording = flag;
	impl->rec = Impl::Record{};
}

bool DeviceWorld::Recording() const
{
	return impl->recording;
}

void* DeviceWorld::Record()
{
	return &impl->rec;
}

struct Batch::Impl
{
	hipStream_t stream = nullptr;
	hipEvent_t ev0 = nullptr, ev1 = nullptr;
	char* staging = nullptr; // pinned
	size_t stagingCapacity = 0;
	DevArray<char> buffer;
	int* queue = nullptr;
	int cuCount = 0;

	size_t used = 0;

	void EnsureStaging( size_t bytes )
	{
		if ( bytes <= stagingCapacity )
		{
			return;
		}
		size_t cap = std::max( bytes, stagingCapacity + stagingCapacity / 2 );
		if ( staging != nullptr )
		{
			HIP_CHECK( hipHostFree( staging ) );
		}
		HIP_CHECK( hipHostMalloc( (void**)&staging, cap, hipHostMallocDefault ) );
		stagingCapacity = cap;
	}

	// Append a block of host data to the staging area (16-byte aligned); returns its offset.
	size_t Put( const void* data, size_t bytes )
	{
		size_t offset = ( used + 15 ) & ~size_t( 15 );
		memcpy( staging + offset, data, bytes );
		used = offset + bytes;
		return offset;
	}
};

Batch::Batch()
{
	impl = new Impl();
	HIP_CHECK( hipStreamCreateWithFlags( &impl->stream, hipStreamNonBlocking ) );
	HIP_CHECK( hipEventCreate( &impl->ev0 ) );
	HIP_CHECK( hipEventCreate( &impl->ev1 ) );
	HIP_CHECK( hipMalloc( &impl->queue, sizeof( int ) ) );
	int device = 0;
	HIP_CHECK( hipGetDevice( &device ) );
	HIP_CHECK( hipDeviceGetAttribute( &impl->cuCount, hipDeviceAttributeMultiprocessorCount, device ) );
}

Batch::~Batch()
{
	(void)hipStreamSynchronize( impl->stream );
	impl->buffer.Free();
	(void)hipFree( impl->queue );
	if ( impl->staging != nullptr )
	{
		(void)hipHostFree( impl->staging );
	}
	(void)hipEventDestroy( impl->ev0 );
	(void)hipEventDestroy( impl->ev1 );
	(void)hipStreamDestroy( impl->stream );
	delete impl;
}

// Per-stage lists gathered from the worlds' records.
template <typename Args>
struct StageList
{
	std::vector<Args> args;
	std::vector<int> lanes;
};

// Upload a lane stage (arguments and block table) into the staging area; returns the offsets.
template <typename Args>
static void StageLane( Batch::Impl* b, const StageList<Args>& list, size_t* argsOffset, size_t* blocksOffset, int* blockCount )
{
	std::vector<int2> blocks;
	for ( size_t w = 0; w < list.args.size(); ++w )
	{
		int n = ( list.lanes[w] + kBlock - 1 ) / kBlock;
		for ( int k = 0; k < n; ++k )
		{
			blocks.push_back( int2{ (int)w, k } );
		}
	}
	*argsOffset = b->Put( list.args.data(), list.args.size() * sizeof( Args ) );
	*blocksOffset = b->Put( blocks.data(), blocks.size() * sizeof( int2 ) );
	*blockCount = (int)blocks.size();
}

// Staging bytes of a lane stage: arguments, block table, alignment slack.
template <typename Args>
static size_t ListBytes( const StageList<Args>& list )
{
	size_t blocks = 0;
	for ( int n : list.lanes )
	{
		blocks += ( n + kBlock - 1 ) / kBlock;
	}
	return list.args.size() * sizeof( Args ) + blocks * sizeof( int2 ) + 32;
}

// Put every world's recorded uploads into the staging area and build the scatter table; the
// worlds' upload records are cleared. Returns the number of entries.
static int StageUploads( Batch::Impl* b, DeviceWorld** worlds, int count, size_t* entriesOffset )
{
	std::vector<ScatterEntry> entries;
	for ( int i = 0; i < count; ++i )
	{
		auto* r = static_cast<DeviceWorld::Impl::Record*>( worlds[i]->Record() );
		if ( r->uploads.empty() )
		{
			continue;
		}
		size_t base = b->Put( r->payload.data(), r->payload.size() );
		for ( const auto& u : r->uploads )
		{
			ScatterEntry e;
			e.dst = static_cast<uint32_t*>( u.dst );
			e.srcOffset = base + u.offset;
			e.words = (uint32_t)( u.bytes / 4 );
			e.pad = 0;
			entries.push_back( e );
		}
		r->uploads.clear();
		r->payload.clear();
	}
	*entriesOffset = b->Put( entries.data(), entries.size() * sizeof( ScatterEntry ) );
	return (int)entries.size();
}

static size_t UploadBytes( DeviceWorld** worlds, int count )
{
	size_t bytes = 0;
	for ( int i = 0; i < count; ++i )
	{
		auto* r = static_cast<DeviceWorld::Impl::Record*>( worlds[i]->Record() );
		bytes += r->payload.size() + r->uploads.size() * sizeof( ScatterEntry ) + 16;
	}
	return bytes + 64;
}

void Batch::Collide( DeviceWorld** worlds, int count )
{
	StageList<CollideArgs> collide;
	StageList<InitArgs> inits;
	std::vector<PackCollideArgs> pack;
	for ( int i = 0; i < count; ++i )
	{
		auto* r = static_cast<DeviceWorld::Impl::Record*>( worlds[i]->Record() );
		if ( r->collide )
		{
			collide.args.push_back( r->collideArgs );
			collide.lanes.push_back( r->collideLanes );
			pack.push_back( r->packCollideArgs );
		}
		if ( r->init )
		{
			inits.args.push_back( r->initArgs );
			inits.lanes.push_back( r->initLanes );
			r->init = false;
		}
	}

	Impl* b = impl;
	b->used = 0;
	b->EnsureStaging( UploadBytes( worlds, count ) + ListBytes( inits ) + ListBytes( collide ) +
					  pack.size() * sizeof( PackCollideArgs ) + 128 );
	size_t entriesOff, initOff, initBlocksOff, argsOff, blocksOff, packOff;
	int initBlocks, blockCount;
	int entryCount = StageUploads( b, worlds, count, &entriesOff );
	StageLane( b, inits, &initOff, &initBlocksOff, &initBlocks );
	StageLane( b, collide, &argsOff, &blocksOff, &blockCount );
	packOff = b->Put( pack.data(), pack.size() * sizeof( PackCollideArgs ) );

	b->buffer.Reserve( b->used, false, b->stream, 0 );
	HIP_CHECK( hipMemcpyAsync( b->buffer.ptr, b->staging, b->used, hipMemcpyHostToDevice, b->stream ) );
	char* base = b->buffer.ptr;
	if ( entryCount > 0 )
	{
		ScatterKernel<<<entryCount, kBlock, 0, b->stream>>>( reinterpret_cast<const ScatterEntry*>( base + entriesOff ), base );
	}
	if ( initBlocks > 0 )
	{
		LaneBatchKernel<InitContactsOp><<<initBlocks, kBlock, 0, b->stream>>>( reinterpret_cast<const InitArgs*>( base + initOff ),
																			   reinterpret_cast<const int2*>( base + initBlocksOff ) );
	}
	if ( blockCount == 0 )
	{
		HIP_CHECK( hipGetLastError() );
		HIP_CHECK( hipStreamSynchronize( b->stream ) );
		return;
	}
	LaneBatchKernel<CollideOp><<<blockCount, kBlock, 0, b->stream>>>( reinterpret_cast<const CollideArgs*>( base + argsOff ),
																	reinterpret_cast<const int2*>( base + blocksOff ) );
	HIP_CHECK( hipGetLastError() );
	PackCollideBatchKernel<<<(int)pack.size(), kPackThreads, 0, b->stream>>>( reinterpret_cast<const PackCollideArgs*>( base + packOff ) );
	HIP_CHECK( hipGetLastError() );
	HIP_CHECK( hipStreamSynchronize( b->stream ) );
}


void Batch::Solve( DeviceWorld** worlds, int count )
{
	StageList<PrepareArgs> prepare;
	StageList<PrepareJointArgs> prepareJoints;
	StageList<TapeRecordArgs> tapeRecords;
	std::vector<FusedArgs> fused;
	std::vector<int> fusedBodies;
	StageList<StoreArgs> store;
	StageList<FinalizeArgs> finalize;
	StageList<ContinuousArgs> continuous;
	std::vector<PackFinalizeArgs> pack;
	double fusedBytes = 0.0;

	for ( int i = 0; i < count; ++i )
	{
		auto* r = static_cast<DeviceWorld::Impl::Record*>( worlds[i]->Record() );
		if ( r->prepare )
		{
			prepare.args.push_back( r->prepareArgs );
			prepare.lanes.push_back( r->prepareLanes );
		}
		if ( r->prepareJoints )
		{
			prepareJoints.args.push_back( r->prepareJointArgs );
			prepareJoints.lanes.push_back( r->prepareJointLanes );
		}
		if ( r->tapeRecord )
		{
			tapeRecords.args.push_back( r->tapeArgs );
			tapeRecords.lanes.push_back( r->tapeLanes );
		}
		if ( r->fused )
		{
			const FusedArgs& fa = r->fusedArgs;
			fused.push_back( fa );
			fusedBodies.push_back( fa.bodyCount );
			double nc = fa.overflowContactStart + fa.overflowContactCount;
			double nj = fa.colorJointStart[fa.activeColorCount] + fa.overflowJointCount;
			fusedBytes += fa.subStepCount * ( kBytesFusedContactSubstep * nc + kBytesFusedBodySubstep * fa.bodyCount + kBytesJoint * nj ) +
						  kBytesFusedBodyStep * fa.bodyCount;
		}
		if ( r->store )
		{
			store.args.push_back( r->storeArgs );
			store.lanes.push_back( r->storeLanes );
		}
		if ( r->finalize )
		{
			finalize.args.push_back( r->finalizeArgs );
			finalize.lanes.push_back( r->finalizeLanes );
		}
		if ( r->continuous )
		{
			continuous.args.push_back( r->continuousArgs );
			continuous.lanes.push_back( r->continuousLanes );
		}
		if ( r->packFinalize )
		{
			pack.push_back( r->packFinalizeArgs );
		}
	}

	Impl* b = impl;
	b->used = 0;
	size_t n = (size_t)count;
	b->EnsureStaging( UploadBytes( worlds, count ) + ListBytes( prepare ) + ListBytes( prepareJoints ) + ListBytes( tapeRecords ) +
					  ListBytes( store ) +
					  ListBytes( finalize ) + ListBytes( continuous ) +
					  n * ( sizeof( FusedArgs ) + sizeof( int ) + sizeof( PackFinalizeArgs ) ) + 256 );
	size_t entriesOff = 0;
	int entryCount = StageUploads( b, worlds, count, &entriesOff );

	size_t prepOff = 0, prepBlocksOff = 0, jointOff = 0, jointBlocksOff = 0, storeOff = 0, storeBlocksOff = 0;
	size_t finOff = 0, finBlocksOff = 0, ccdOff = 0, ccdBlocksOff = 0, fusedOff = 0, orderOff = 0, packOff = 0;
	int prepBlocks = 0, jointBlocks = 0, storeBlocks = 0, finBlocks = 0, ccdBlocks = 0;
	size_t tapeOff = 0, tapeBlocksOff = 0;
	int tapeBlocks = 0;
	StageLane( b, prepare, &prepOff, &prepBlocksOff, &prepBlocks );
	StageLane( b, prepareJoints, &jointOff, &jointBlocksOff, &jointBlocks );
	StageLane( b, tapeRecords, &tapeOff, &tapeBlocksOff, &tapeBlocks );
	StageLane( b, store, &storeOff, &storeBlocksOff, &storeBlocks );
	StageLane( b, finalize, &finOff, &finBlocksOff, &finBlocks );
	StageLane( b, continuous, &ccdOff, &ccdBlocksOff, &ccdBlocks );

	// largest worlds first, so the persistent solver's queue ends with the short jobs
	std::vector<int> order( fused.size() );
	for ( size_t i = 0; i < order.size(); ++i )
	{
		order[i] = (int)i;
	}
	std::stable_sort( order.begin(), order.end(), [&]( int x, int y ) { return fusedBodies[x] > fusedBodies[y]; } );
	fusedOff = b->Put( fused.data(), fused.size() * sizeof( FusedArgs ) );
	orderOff = b->Put( order.data(), order.size() * sizeof( int ) );
	packOff = b->Put( pack.data(), pack.size() * sizeof( PackFinalizeArgs ) );

	b->buffer.Reserve( b->used, false, b->stream, 0 );
	hipStream_t s = b->stream;
	HIP_CHECK( hipMemcpyAsync( b->buffer.ptr, b->staging, b->used, hipMemcpyHostToDevice, s ) );
	char* base = b->buffer.ptr;

	if ( entryCount > 0 )
	{
		ScatterKernel<<<entryCount, kBlock, 0, s>>>( reinterpret_cast<const ScatterEntry*>( base + entriesOff ), base );
	}
	if ( prepBlocks > 0 )
	{
		LaneBatchKernel<PrepareContactsOp><<<prepBlocks, kBlock, 0, s>>>( reinterpret_cast<const PrepareArgs*>( base + prepOff ),
																	   reinterpret_cast<const int2*>( base + prepBlocksOff ) );
	}
	if ( jointBlocks > 0 )
	{
		LaneBatchKernel<PrepareJointsOp><<<jointBlocks, kBlock, 0, s>>>( reinterpret_cast<const PrepareJointArgs*>( base + jointOff ),
																	  reinterpret_cast<const int2*>( base + jointBlocksOff ) );
	}
	if ( tapeBlocks > 0 )
	{
		LaneBatchKernel<TapeRecordOp><<<tapeBlocks, kBlock, 0, s>>>( reinterpret_cast<const TapeRecordArgs*>( base + tapeOff ),
																  reinterpret_cast<const int2*>( base + tapeBlocksOff ) );
	}

	for ( int k = 0; k < kBatchBuckets; ++k )
	{
		lastBucketWorlds[k] = 0;
		lastBucketGrid[k] = 0;
		lastBucketMs[k] = 0.0f;
	}
	if ( fused.empty() == false )
	{
		// One persistent launch over every world, largest first. Its workgroup size is the tuned
		// table's for the bucket with the most worlds (the worlds of one launch share one size; a
		// launch per bucket, tried side by side on separate streams, lost to the single queue's load
		// balance); each world's warm-start form is its own bucket's. Neither changes any result.
		const int total = (int)fused.size();
		int bucketCount[kBatchBuckets] = {};
		for ( int i = 0; i < total; ++i )
		{
			bucketCount[BatchBucketOf( fusedBodies[i] )] += 1;
		}
		int most = 0;
		for ( int k = 1; k < kBatchBuckets; ++k )
		{
			most = bucketCount[k] > bucketCount[most] ? k : most;
		}
		const int threads = BatchTuningFor( most ).threads;
		int lds = FusedLdsBytes( fusedBodies[order[0]] );
		int perCu = 0;
		HIP_CHECK( hipOccupancyMaxActiveBlocksPerMultiprocessor( &perCu, SolveFusedBatchKernel, threads, lds ) );
		int grid = std::max( 1, std::min( total, perCu * b->cuCount ) );
		HIP_CHECK( hipMemsetAsync( b->queue, 0, sizeof( int ), s ) );
		// the events bracket the kernel's execution (recorded by the dispatch itself)
		hipExtLaunchKernelGGL( SolveFusedBatchKernel, dim3( grid ), dim3( threads ), (uint32_t)lds, s, b->ev0, b->ev1, 0,
							   reinterpret_cast<const FusedArgs*>( base + fusedOff ), reinterpret_cast<const int*>( base + orderOff ),
							   total, b->queue );
		HIP_CHECK( hipGetLastError() );
		for ( int k = 0; k < kBatchBuckets; ++k )
		{
			lastBucketWorlds[k] = bucketCount[k];
			lastBucketThreads[k] = bucketCount[k] > 0 ? threads : 0;
			lastBucketWarmByBody[k] = bucketCount[k] > 0 ? BatchTuningFor( k ).warmByBody : 0;
			lastBucketGrid[k] = bucketCount[k] > 0 ? grid : 0;
		}
		lastFusedBytes = fusedBytes;
	}

	if ( storeBlocks > 0 )
	{
		LaneBatchKernel<StoreImpulsesOp><<<storeBlocks, kBlock, 0, s>>>( reinterpret_cast<const StoreArgs*>( base + storeOff ),
																	  reinterpret_cast<const int2*>( base + storeBlocksOff ) );
	}
	if ( finBlocks > 0 )
	{
		LaneBatchKernel<FinalizeBodiesOp><<<finBlocks, kBlock, 0, s>>>( reinterpret_cast<const FinalizeArgs*>( base + finOff ),
																	 reinterpret_cast<const int2*>( base + finBlocksOff ) );
	}
	if ( ccdBlocks > 0 )
	{
		LaneBatchKernel<ContinuousOp><<<ccdBlocks, kBlock, 0, s>>>( reinterpret_cast<const ContinuousArgs*>( base + ccdOff ),
																 reinterpret_cast<const int2*>( base + ccdBlocksOff ) );
	}
	if ( pack.empty() == false )
	{
		PackFinalizeBatchKernel<<<(int)pack.size(), kPackThreads, 0, s>>>( reinterpret_cast<const PackFinalizeArgs*>( base + packOff ) );
	}
	HIP_CHECK( hipGetLastError() );
	HIP_CHECK( hipStreamSynchronize( s ) );

	// worlds beyond the persistent solver: their own solve, now that the batch has applied their
	// uploads (each one's StepEnd waits for it)
	for ( int i = 0; i < count; ++i )
	{
		worlds[i]->LaunchDeferredSolve();
	}

	if ( fused.empty() == false )
	{
		float ms = 0.0f;
		HIP_CHECK( hipEventElapsedTime( &ms, b->ev0, b->ev1 ) );
		lastFusedMs = ms;
		for ( int k = 0; k < kBatchBuckets; ++k )
		{
			lastBucketMs[k] = lastBucketWorlds[k] > 0 ? ms : 0.0f;
		}
	}
}

} // namespace b2g
