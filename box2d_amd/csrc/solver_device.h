// SPDX-FileCopyrightText: 2022 Erin Catto (Box2D; the contact solver arithmetic below restates its
// src/contact_solver.c and src/solver.c in the same operation order)
// SPDX-License-Identifier: MIT
// GPU solver arithmetic for one contact or joint constraint.
//
// These functions are the per-lane bodies of the constraint kernels. The colored path reproduces
// the reference's wide (SIMD lane) arithmetic, including its operation order, its a >= b / a <= b
// min-max semantics and the fact that a one-point manifold still runs the second point with zeros
// (reference src/contact_solver.c:936-1063, 1811-2228). The overflow path reproduces the scalar
// overflow solver (src/contact_solver.c:24-545), whose expressions differ from the wide ones.
// Joints are in joints_device.h.
#pragma once

#include "layout.h"

namespace b2g
{

#define B2G_DEVI __device__ __forceinline__

// The constraint fields live in one allocation, field-major: field k of constraint i is at
// base[k * stride + i]. f[k] yields the field's array, so lanes read c.f[k][i].
struct FieldTable
{
	float* base;
	int stride;

	__host__ __device__ float* operator[]( int field ) const
	{
		return base + (size_t)field * stride;
	}
};

struct ConstraintSoA
{
	FieldTable f;
	int* bodyA; // body ids (GlobalBodies view)
	int* bodyB;
	int* contactId;
	int* pointCount;
	int* slotA; // LDS slots of the fused solver (LdsBodies view)
	int* slotB;
};

// LDS body slots of the fused solver: every static body reads the identity state kept in slot 0,
// awake bodies occupy slots 1..count (at most kFusedBodyCapacity).
constexpr int kFusedBodyCapacity = 5400;
constexpr int kStaticSlot = 0;

B2G_DEVI float MaxW( float a, float b )
{
	return a >= b ? a : b;
}

B2G_DEVI float MinW( float a, float b )
{
	return a <= b ? a : b;
}

B2G_DEVI Softness MakeSoft( float hertz, float zeta, float h )
{
	if ( hertz == 0.0f )
	{
		return Softness{ 0.0f, 0.0f, 0.0f };
	}
	float omega = 2.0f * B2_PI * hertz;
	float a1 = 2.0f * zeta + h * omega;
	float a2 = h * omega * a1;
	float a3 = 1.0f / ( 1.0f + a2 );
	return Softness{ omega / a1, a2 * a3, a3 };
}

// Deterministic arctangent (reference src/math_functions.c:96-136).
B2G_DEVI float Atan2( float y, float x )
{
	if ( x == 0.0f && y == 0.0f )
	{
		return 0.0f;
	}
	float ax = b2AbsFloat( x );
	float ay = b2AbsFloat( y );
	float mx = b2MaxFloat( ay, ax );
	float mn = b2MinFloat( ay, ax );
	float a = mn / mx;
	float s = a * a;
	float c = s * a;
	float q = s * s;
	float r = 0.024840285f * q + 0.18681418f;
	float t = -0.094097948f * q - 0.33213072f;
	r = r * s + t;
	r = r * c + a;
	if ( ay > ax )
	{
		r = 1.57079637f - r;
	}
	if ( x < 0 )
	{
		r = 3.14159274f - r;
	}
	if ( y < 0 )
	{
		r = -r;
	}
	return r;
}

struct LaneBody
{
	float vx, vy, w;
	uint32_t flags;
	float dpx, dpy, dqc, dqs;
};

// Body state access of the solver functions. GlobalBodies reads and writes the id-indexed state
// array in device memory; LdsBodies reads and writes a world's awake bodies staged in LDS by
// solver-set index (the fused solver kernel). Constraint and joint body indices must match the
// view: body ids for GlobalBodies, LDS slots for LdsBodies.
struct GlobalBodies
{
	BodyState* states;

	B2G_DEVI LaneBody Load( int index ) const
	{
		const float4* p = reinterpret_cast<const float4*>( states + index );
		float4 a = p[0];
		float4 b = p[1];
		LaneBody r;
		r.vx = a.x;
		r.vy = a.y;
		r.w = a.z;
		r.flags = __float_as_uint( a.w );
		r.dpx = b.x;
		r.dpy = b.y;
		r.dqc = b.z;
		r.dqs = b.w;
		return r;
	}

	// Delta position and rotation only (the fields no solver stage writes).
	B2G_DEVI void LoadDelta( int index, b2Vec2& dp, b2Rot& dq ) const
	{
		float4 b = reinterpret_cast<const float4*>( states + index )[1];
		dp = { b.x, b.y };
		dq = { b.z, b.w };
	}

	// Velocities are written back for dynamic bodies only (kinematic and static states are read-only).
	B2G_DEVI void Store( int index, const LaneBody& b ) const
	{
		if ( b.flags & kDynamicFlag )
		{
			BodyState* s = states + index;
			s->linearVelocity.x = b.vx;
			s->linearVelocity.y = b.vy;
			s->angularVelocity = b.w;
		}
	}

	B2G_DEVI static int JointA( const JointSim* j )
	{
		return j->indexA;
	}
	B2G_DEVI static int JointB( const JointSim* j )
	{
		return j->indexB;
	}
};

// GlobalBodies for solvers whose workgroups exchange body states inside one launch (the XCD
// solver): loads are agent-scope relaxed atomic loads, which skip the CU's L1, where another
// CU's earlier copy of the line may still sit, and read the L2 that every writer's stores reach.
// Stores are plain: the L1 writes through to the L2.
struct CoherentBodies
{
	BodyState* states;

	B2G_DEVI static uint64_t Word( const BodyState* s, int k )
	{
		return __hip_atomic_load( reinterpret_cast<const uint64_t*>( s ) + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
	}

	B2G_DEVI static float Lo( uint64_t u )
	{
		return __uint_as_float( (uint32_t)u );
	}

	B2G_DEVI static float Hi( uint64_t u )
	{
		return __uint_as_float( (uint32_t)( u >> 32 ) );
	}

	B2G_DEVI LaneBody Load( int index ) const
	{
		const BodyState* s = states + index;
		uint64_t a = Word( s, 0 ), b = Word( s, 1 ), c = Word( s, 2 ), d = Word( s, 3 );
		LaneBody r;
		r.vx = Lo( a );
		r.vy = Hi( a );
		r.w = Lo( b );
		r.flags = (uint32_t)( b >> 32 );
		r.dpx = Lo( c );
		r.dpy = Hi( c );
		r.dqc = Lo( d );
		r.dqs = Hi( d );
		return r;
	}

	B2G_DEVI void LoadDelta( int index, b2Vec2& dp, b2Rot& dq ) const
	{
		const BodyState* s = states + index;
		uint64_t c = Word( s, 2 ), d = Word( s, 3 );
		dp = { Lo( c ), Hi( c ) };
		dq = { Lo( d ), Hi( d ) };
	}

	B2G_DEVI void Store( int index, const LaneBody& b ) const
	{
		if ( b.flags & kDynamicFlag )
		{
			BodyState* s = states + index;
			s->linearVelocity.x = b.vx;
			s->linearVelocity.y = b.vy;
			s->angularVelocity = b.w;
		}
	}

	// The whole state of a body (for the per-body integration phases).
	B2G_DEVI BodyState LoadState( int index ) const
	{
		LaneBody l = Load( index );
		BodyState st;
		st.linearVelocity = { l.vx, l.vy };
		st.angularVelocity = l.w;
		st.flags = l.flags;
		st.deltaPosition = { l.dpx, l.dpy };
		st.deltaRotation = { l.dqc, l.dqs };
		return st;
	}

	B2G_DEVI static int JointA( const JointSim* j )
	{
		return j->indexA;
	}
	B2G_DEVI static int JointB( const JointSim* j )
	{
		return j->indexB;
	}
};

struct LdsBodies
{
	float* vx;
	float* vy;
	float* w;
	float* dpx;
	float* dpy;
	float* dqc;
	float* dqs;
	uint16_t* flags;

	B2G_DEVI LaneBody Load( int index ) const
	{
		LaneBody r;
		r.vx = vx[index];
		r.vy = vy[index];
		r.w = w[index];
		r.flags = flags[index];
		r.dpx = dpx[index];
		r.dpy = dpy[index];
		r.dqc = dqc[index];
		r.dqs = dqs[index];
		return r;
	}

	B2G_DEVI void LoadDelta( int index, b2Vec2& dp, b2Rot& dq ) const
	{
		dp = { dpx[index], dpy[index] };
		dq = { dqc[index], dqs[index] };
	}

	B2G_DEVI void Store( int index, const LaneBody& b ) const
	{
		if ( b.flags & kDynamicFlag )
		{
			vx[index] = b.vx;
			vy[index] = b.vy;
			w[index] = b.w;
		}
	}

	B2G_DEVI static int JointA( const JointSim* j )
	{
		return j->slotA;
	}
	B2G_DEVI static int JointB( const JointSim* j )
	{
		return j->slotB;
	}
};

// ---------------------------------------------------------------------------------------------
// Body integration

// Velocity integration of one body (reference src/solver.c:66-112).
B2G_DEVI void IntegrateVelocity( const BodySim& sim, BodyState* state, const StepParams& p )
{
	b2Vec2 v = state->linearVelocity;
	float w = state->angularVelocity;
	float h = p.h;

	float linearDamping = 1.0f / ( 1.0f + h * sim.linearDamping );
	float angularDamping = 1.0f / ( 1.0f + h * sim.angularDamping );
	float gravityScale = sim.invMass > 0.0f ? sim.gravityScale : 0.0f;

	b2Vec2 linearVelocityDelta = b2Add( b2MulSV( h * sim.invMass, sim.force ), b2MulSV( h * gravityScale, p.gravity ) );
	float angularVelocityDelta = h * sim.invInertia * sim.torque;

	v = b2MulAdd( linearVelocityDelta, linearDamping, v );
	w = angularVelocityDelta + angularDamping * w;

	state->linearVelocity = v;
	state->angularVelocity = w;
}

// Position integration of one body: locks, speed caps, delta position and rotation
// (reference src/solver.c:114-162).
B2G_DEVI void IntegratePosition( BodyState* state, const StepParams& p )
{
	float h = p.h;
	float maxLinearSpeed = p.maxLinearVelocity;
	float maxAngularSpeed = ( 0.25f * B2_PI ) * p.inv_dt;
	float maxLinearSpeedSquared = maxLinearSpeed * maxLinearSpeed;
	float maxAngularSpeedSquared = maxAngularSpeed * maxAngularSpeed;

	b2Vec2 v = state->linearVelocity;
	float w = state->angularVelocity;
	uint32_t flags = state->flags;

	v.x = ( flags & kLockLinearX ) ? 0.0f : v.x;
	v.y = ( flags & kLockLinearY ) ? 0.0f : v.y;
	w = ( flags & kLockAngularZ ) ? 0.0f : w;

	if ( b2Dot( v, v ) > maxLinearSpeedSquared )
	{
		float ratio = maxLinearSpeed / b2Length( v );
		v = b2MulSV( ratio, v );
		flags |= kIsSpeedCapped;
	}

	if ( w * w > maxAngularSpeedSquared && ( flags & kAllowFastRotation ) == 0 )
	{
		float ratio = maxAngularSpeed / b2AbsFloat( w );
		w *= ratio;
		flags |= kIsSpeedCapped;
	}

	state->flags = flags;
	state->linearVelocity = v;
	state->angularVelocity = w;
	state->deltaPosition = b2MulAdd( state->deltaPosition, h, v );
	state->deltaRotation = b2IntegrateRotation( state->deltaRotation, h * w );
}

// ---------------------------------------------------------------------------------------------
// Colored (wide-lane) contact arithmetic

template <typename Bodies>
B2G_DEVI void WarmStartContact( const ConstraintSoA& c, int i, const Bodies& B )
{
	int ia = c.bodyA[i];
	int ib = c.bodyB[i];
	LaneBody bA = B.Load( ia );
	LaneBody bB = B.Load( ib );

	float mA = c.f[kcInvMassA][i], mB = c.f[kcInvMassB][i];
	float iA = c.f[kcInvIA][i], iB = c.f[kcInvIB][i];
	float nX = c.f[kcNormalX][i], nY = c.f[kcNormalY][i];
	float tX = nY;
	float tY = 0.0f - nX;

	{
		float rAx = c.f[kcAnchorA1X][i], rAy = c.f[kcAnchorA1Y][i];
		float rBx = c.f[kcAnchorB1X][i], rBy = c.f[kcAnchorB1Y][i];
		float ni = c.f[kcNormalImpulse1][i], ti = c.f[kcTangentImpulse1][i];
		float Px = ni * nX + ti * tX;
		float Py = ni * nY + ti * tY;
		bA.w = bA.w - iA * ( rAx * Py - rAy * Px );
		bA.vx = bA.vx - mA * Px;
		bA.vy = bA.vy - mA * Py;
		bB.w = bB.w + iB * ( rBx * Py - rBy * Px );
		bB.vx = bB.vx + mB * Px;
		bB.vy = bB.vy + mB * Py;
		c.f[kcTotalNormalImpulse1][i] = c.f[kcTotalNormalImpulse1][i] + ni;
	}

	{
		float rAx = c.f[kcAnchorA2X][i], rAy = c.f[kcAnchorA2Y][i];
		float rBx = c.f[kcAnchorB2X][i], rBy = c.f[kcAnchorB2Y][i];
		float ni = c.f[kcNormalImpulse2][i], ti = c.f[kcTangentImpulse2][i];
		float Px = ni * nX + ti * tX;
		float Py = ni * nY + ti * tY;
		bA.w = bA.w - iA * ( rAx * Py - rAy * Px );
		bA.vx = bA.vx - mA * Px;
		bA.vy = bA.vy - mA * Py;
		bB.w = bB.w + iB * ( rBx * Py - rBy * Px );
		bB.vx = bB.vx + mB * Px;
		bB.vy = bB.vy + mB * Py;
		c.f[kcTotalNormalImpulse2][i] = c.f[kcTotalNormalImpulse2][i] + ni;
	}

	float rolling = c.f[kcRollingImpulse][i];
	bA.w = bA.w - iA * rolling;
	bB.w = bB.w + iB * rolling;

	B.Store( ia, bA );
	B.Store( ib, bB );
}

// One normal point of the wide solve. Returns the new accumulated impulse for the rolling term.
B2G_DEVI float SolveNormalPoint( const ConstraintSoA& c, int i, LaneBody& bA, LaneBody& bB, float mA, float mB, float iA,
								 float iB, float nX, float nY, float dpX, float dpY, float biasRate, float massScale,
								 float impulseScale, float inv_h, float contactSpeed, int fAx, int fAy, int fBx, int fBy,
								 int fBase, int fNormalImpulse, int fTotal, int fNormalMass )
{
	float rAx = c.f[fAx][i], rAy = c.f[fAy][i];
	float rBx = c.f[fBx][i], rBy = c.f[fBy][i];

	// moving anchors for the current separation
	float rsAx = bA.dqc * rAx - bA.dqs * rAy;
	float rsAy = bA.dqs * rAx + bA.dqc * rAy;
	float rsBx = bB.dqc * rBx - bB.dqs * rBy;
	float rsBy = bB.dqs * rBx + bB.dqc * rBy;

	float dsx = dpX + ( rsBx - rsAx );
	float dsy = dpY + ( rsBy - rsAy );
	float s = ( nX * dsx + nY * dsy ) + c.f[fBase][i];

	bool mask = s > 0.0f;
	float specBias = s * inv_h;
	float softBias = MaxW( biasRate * s, contactSpeed );
	float bias = mask ? specBias : softBias;
	float pointMassScale = mask ? 1.0f : massScale;
	float pointImpulseScale = mask ? 0.0f : impulseScale;

	float dvx = ( bB.vx - bB.w * rBy ) - ( bA.vx - bA.w * rAy );
	float dvy = ( bB.vy + bB.w * rBx ) - ( bA.vy + bA.w * rAx );
	float vn = dvx * nX + dvy * nY;

	float normalImpulse = c.f[fNormalImpulse][i];
	float negImpulse = c.f[fNormalMass][i] * ( pointMassScale * vn + bias ) + pointImpulseScale * normalImpulse;
	float newImpulse = MaxW( normalImpulse - negImpulse, 0.0f );
	float impulse = newImpulse - normalImpulse;
	c.f[fNormalImpulse][i] = newImpulse;
	c.f[fTotal][i] = c.f[fTotal][i] + impulse;

	float Px = impulse * nX;
	float Py = impulse * nY;
	bA.vx = bA.vx - mA * Px;
	bA.vy = bA.vy - mA * Py;
	bA.w = bA.w - iA * ( rAx * Py - rAy * Px );
	bB.vx = bB.vx + mB * Px;
	bB.vy = bB.vy + mB * Py;
	bB.w = bB.w + iB * ( rBx * Py - rBy * Px );
	return newImpulse;
}

B2G_DEVI void FrictionPoint( const ConstraintSoA& c, int i, LaneBody& bA, LaneBody& bB, float mA, float mB, float iA, float iB,
							 float tX, float tY, float friction, float tangentSpeed, int fAx, int fAy, int fBx, int fBy,
							 int fNormalImpulse, int fTangentImpulse, int fTangentMass )
{
	float rAx = c.f[fAx][i], rAy = c.f[fAy][i];
	float rBx = c.f[fBx][i], rBy = c.f[fBy][i];

	float dvx = ( bB.vx - bB.w * rBy ) - ( bA.vx - bA.w * rAy );
	float dvy = ( bB.vy + bB.w * rBx ) - ( bA.vy + bA.w * rAx );
	float vt = dvx * tX + dvy * tY;
	vt = vt - tangentSpeed;

	float negImpulse = c.f[fTangentMass][i] * vt;
	float maxFriction = friction * c.f[fNormalImpulse][i];
	float tangentImpulse = c.f[fTangentImpulse][i];
	float newImpulse = tangentImpulse - negImpulse;
	newImpulse = MaxW( 0.0f - maxFriction, MinW( newImpulse, maxFriction ) );
	float impulse = newImpulse - tangentImpulse;
	c.f[fTangentImpulse][i] = newImpulse;

	float Px = impulse * tX;
	float Py = impulse * tY;
	bA.vx = bA.vx - mA * Px;
	bA.vy = bA.vy - mA * Py;
	bA.w = bA.w - iA * ( rAx * Py - rAy * Px );
	bB.vx = bB.vx + mB * Px;
	bB.vy = bB.vy + mB * Py;
	bB.w = bB.w + iB * ( rBx * Py - rBy * Px );
}

// laneGroupRolling: true when any constraint of this constraint's 4-lane group has rolling
// resistance (the wide solver tests the whole group).
template <typename Bodies>
B2G_DEVI void SolveContact( const ConstraintSoA& c, int i, const Bodies& B, bool useBias, float inv_h, float contactSpeed,
							bool laneGroupRolling )
{
	int ia = c.bodyA[i];
	int ib = c.bodyB[i];
	LaneBody bA = B.Load( ia );
	LaneBody bB = B.Load( ib );

	float mA = c.f[kcInvMassA][i], mB = c.f[kcInvMassB][i];
	float iA = c.f[kcInvIA][i], iB = c.f[kcInvIB][i];
	float nX = c.f[kcNormalX][i], nY = c.f[kcNormalY][i];

	float biasRate, massScale, impulseScale;
	if ( useBias )
	{
		biasRate = c.f[kcMassScale][i] * c.f[kcBiasRate][i];
		massScale = c.f[kcMassScale][i];
		impulseScale = c.f[kcImpulseScale][i];
	}
	else
	{
		biasRate = 0.0f;
		massScale = 1.0f;
		impulseScale = 0.0f;
	}

	float dpX = bB.dpx - bA.dpx;
	float dpY = bB.dpy - bA.dpy;

	float totalNormalImpulse = 0.0f;
	float n1 = SolveNormalPoint( c, i, bA, bB, mA, mB, iA, iB, nX, nY, dpX, dpY, biasRate, massScale, impulseScale, inv_h,
								 contactSpeed, kcAnchorA1X, kcAnchorA1Y, kcAnchorB1X, kcAnchorB1Y, kcBaseSeparation1,
								 kcNormalImpulse1, kcTotalNormalImpulse1, kcNormalMass1 );
	totalNormalImpulse = totalNormalImpulse + n1;
	float n2 = SolveNormalPoint( c, i, bA, bB, mA, mB, iA, iB, nX, nY, dpX, dpY, biasRate, massScale, impulseScale, inv_h,
								 contactSpeed, kcAnchorA2X, kcAnchorA2Y, kcAnchorB2X, kcAnchorB2Y, kcBaseSeparation2,
								 kcNormalImpulse2, kcTotalNormalImpulse2, kcNormalMass2 );
	totalNormalImpulse = totalNormalImpulse + n2;

	if ( useBias == false )
	{
		if ( laneGroupRolling )
		{
			float deltaLambda = c.f[kcRollingMass][i] * ( bA.w - bB.w );
			float lambda = c.f[kcRollingImpulse][i];
			float maxLambda = c.f[kcRollingResistance][i] * totalNormalImpulse;
			float rolling = b2ClampFloat( lambda + deltaLambda, -maxLambda, maxLambda );
			c.f[kcRollingImpulse][i] = rolling;
			deltaLambda = rolling - lambda;
			bA.w = bA.w - iA * deltaLambda;
			bB.w = bB.w + iB * deltaLambda;
		}

		float tX = nY;
		float tY = 0.0f - nX;
		float friction = c.f[kcFriction][i];
		float tangentSpeed = c.f[kcTangentSpeed][i];
		FrictionPoint( c, i, bA, bB, mA, mB, iA, iB, tX, tY, friction, tangentSpeed, kcAnchorA1X, kcAnchorA1Y, kcAnchorB1X,
					   kcAnchorB1Y, kcNormalImpulse1, kcTangentImpulse1, kcTangentMass1 );
		FrictionPoint( c, i, bA, bB, mA, mB, iA, iB, tX, tY, friction, tangentSpeed, kcAnchorA2X, kcAnchorA2Y, kcAnchorB2X,
					   kcAnchorB2Y, kcNormalImpulse2, kcTangentImpulse2, kcTangentMass2 );
	}

	B.Store( ia, bA );
	B.Store( ib, bB );
}

B2G_DEVI void RestitutionPoint( const ConstraintSoA& c, int i, LaneBody& bA, LaneBody& bB, float mA, float mB, float iA,
								float iB, float nX, float nY, float threshold, float restitution, bool restitutionMask,
								int fAx, int fAy, int fBx, int fBy, int fRelVel, int fTotal, int fNormalMass,
								int fNormalImpulse )
{
	float relativeVelocity = c.f[fRelVel][i];
	bool mask1 = ( relativeVelocity + threshold ) > 0.0f;
	bool mask2 = c.f[fTotal][i] == 0.0f;
	bool mask = mask1 || mask2 || restitutionMask;
	float mass = mask ? 0.0f : c.f[fNormalMass][i];

	float rAx = c.f[fAx][i], rAy = c.f[fAy][i];
	float rBx = c.f[fBx][i], rBy = c.f[fBy][i];

	float dvx = ( bB.vx - bB.w * rBy ) - ( bA.vx - bA.w * rAy );
	float dvy = ( bB.vy + bB.w * rBx ) - ( bA.vy + bA.w * rAx );
	float vn = dvx * nX + dvy * nY;

	float negImpulse = mass * ( vn + restitution * relativeVelocity );
	float normalImpulse = c.f[fNormalImpulse][i];
	float newImpulse = MaxW( normalImpulse - negImpulse, 0.0f );
	float deltaImpulse = newImpulse - normalImpulse;
	c.f[fNormalImpulse][i] = newImpulse;
	c.f[fTotal][i] = c.f[fTotal][i] + deltaImpulse;

	float Px = deltaImpulse * nX;
	float Py = deltaImpulse * nY;
	bA.vx = bA.vx - mA * Px;
	bA.vy = bA.vy - mA * Py;
	bA.w = bA.w - iA * ( rAx * Py - rAy * Px );
	bB.vx = bB.vx + mB * Px;
	bB.vy = bB.vy + mB * Py;
	bB.w = bB.w + iB * ( rBx * Py - rBy * Px );
}

template <typename Bodies>
B2G_DEVI void RestitutionContact( const ConstraintSoA& c, int i, const Bodies& B, float threshold )
{
	int ia = c.bodyA[i];
	int ib = c.bodyB[i];
	LaneBody bA = B.Load( ia );
	LaneBody bB = B.Load( ib );

	float mA = c.f[kcInvMassA][i], mB = c.f[kcInvMassB][i];
	float iA = c.f[kcInvIA][i], iB = c.f[kcInvIB][i];
	float nX = c.f[kcNormalX][i], nY = c.f[kcNormalY][i];
	float restitution = c.f[kcRestitution][i];
	bool restitutionMask = restitution == 0.0f;

	RestitutionPoint( c, i, bA, bB, mA, mB, iA, iB, nX, nY, threshold, restitution, restitutionMask, kcAnchorA1X, kcAnchorA1Y,
					  kcAnchorB1X, kcAnchorB1Y, kcRelativeVelocity1, kcTotalNormalImpulse1, kcNormalMass1, kcNormalImpulse1 );
	RestitutionPoint( c, i, bA, bB, mA, mB, iA, iB, nX, nY, threshold, restitution, restitutionMask, kcAnchorA2X, kcAnchorA2Y,
					  kcAnchorB2X, kcAnchorB2Y, kcRelativeVelocity2, kcTotalNormalImpulse2, kcNormalMass2, kcNormalImpulse2 );

	B.Store( ia, bA );
	B.Store( ib, bB );
}

// ---------------------------------------------------------------------------------------------
// Overflow (scalar) contact arithmetic, solved one constraint after another by a single lane.

B2G_DEVI BodyState ToState( const LaneBody& b )
{
	BodyState s;
	s.linearVelocity = { b.vx, b.vy };
	s.angularVelocity = b.w;
	s.flags = b.flags;
	s.deltaPosition = { b.dpx, b.dpy };
	s.deltaRotation = { b.dqc, b.dqs };
	return s;
}

B2G_DEVI LaneBody ToLane( const BodyState& s )
{
	LaneBody b;
	b.vx = s.linearVelocity.x;
	b.vy = s.linearVelocity.y;
	b.w = s.angularVelocity;
	b.flags = s.flags;
	b.dpx = s.deltaPosition.x;
	b.dpy = s.deltaPosition.y;
	b.dqc = s.deltaRotation.c;
	b.dqs = s.deltaRotation.s;
	return b;
}

struct OverflowVel
{
	float vx, vy, w;
};

template <typename Bodies>
B2G_DEVI void OverflowWarmStart( const ConstraintSoA& c, int i, const Bodies& B )
{
	int ia = c.bodyA[i];
	int ib = c.bodyB[i];
	BodyState sA = ToState( B.Load( ia ) );
	BodyState* stateA = &sA;
	BodyState sB = ToState( B.Load( ib ) );
	BodyState* stateB = &sB;

	b2Vec2 vA = stateA->linearVelocity;
	float wA = stateA->angularVelocity;
	b2Vec2 vB = stateB->linearVelocity;
	float wB = stateB->angularVelocity;

	float mA = c.f[kcInvMassA][i], mB = c.f[kcInvMassB][i];
	float iA = c.f[kcInvIA][i], iB = c.f[kcInvIB][i];
	b2Vec2 normal = { c.f[kcNormalX][i], c.f[kcNormalY][i] };
	b2Vec2 tangent = b2RightPerp( normal );
	int pointCount = c.pointCount[i];

	for ( int j = 0; j < pointCount; ++j )
	{
		int o = j == 0 ? 0 : ( kcAnchorA2X - kcAnchorA1X );
		b2Vec2 rA = { c.f[kcAnchorA1X + o][i], c.f[kcAnchorA1Y + o][i] };
		b2Vec2 rB = { c.f[kcAnchorB1X + o][i], c.f[kcAnchorB1Y + o][i] };
		float ni = c.f[kcNormalImpulse1 + o][i];
		float ti = c.f[kcTangentImpulse1 + o][i];
		b2Vec2 P = b2Add( b2MulSV( ni, normal ), b2MulSV( ti, tangent ) );
		c.f[kcTotalNormalImpulse1 + o][i] = c.f[kcTotalNormalImpulse1 + o][i] + ni;
		wA -= iA * b2Cross( rA, P );
		vA = b2MulAdd( vA, -mA, P );
		wB += iB * b2Cross( rB, P );
		vB = b2MulAdd( vB, mB, P );
	}

	float rolling = c.f[kcRollingImpulse][i];
	wA -= iA * rolling;
	wB += iB * rolling;

	if ( stateA->flags & kDynamicFlag )
	{
		stateA->linearVelocity = vA;
		stateA->angularVelocity = wA;
	}
	if ( stateB->flags & kDynamicFlag )
	{
		stateB->linearVelocity = vB;
		stateB->angularVelocity = wB;
	}

	B.Store( ia, ToLane( sA ) );
	B.Store( ib, ToLane( sB ) );
}

template <typename Bodies>
B2G_DEVI void OverflowSolve( const ConstraintSoA& c, int i, const Bodies& B, bool useBias, float inv_h, float contactSpeed )
{
	float mA = c.f[kcInvMassA][i], mB = c.f[kcInvMassB][i];
	float iA = c.f[kcInvIA][i], iB = c.f[kcInvIB][i];

	int ia = c.bodyA[i];
	int ib = c.bodyB[i];
	BodyState sA = ToState( B.Load( ia ) );
	BodyState* stateA = &sA;
	BodyState sB = ToState( B.Load( ib ) );
	BodyState* stateB = &sB;

	b2Vec2 vA = stateA->linearVelocity;
	float wA = stateA->angularVelocity;
	b2Rot dqA = stateA->deltaRotation;
	b2Vec2 vB = stateB->linearVelocity;
	float wB = stateB->angularVelocity;
	b2Rot dqB = stateB->deltaRotation;

	b2Vec2 dp = b2Sub( stateB->deltaPosition, stateA->deltaPosition );

	b2Vec2 normal = { c.f[kcNormalX][i], c.f[kcNormalY][i] };
	b2Vec2 tangent = b2RightPerp( normal );
	float friction = c.f[kcFriction][i];
	Softness softness = { c.f[kcBiasRate][i], c.f[kcMassScale][i], c.f[kcImpulseScale][i] };
	int pointCount = c.pointCount[i];
	float totalNormalImpulse = 0.0f;

	for ( int j = 0; j < pointCount; ++j )
	{
		int o = j == 0 ? 0 : ( kcAnchorA2X - kcAnchorA1X );
		b2Vec2 rA = { c.f[kcAnchorA1X + o][i], c.f[kcAnchorA1Y + o][i] };
		b2Vec2 rB = { c.f[kcAnchorB1X + o][i], c.f[kcAnchorB1Y + o][i] };

		b2Vec2 ds = b2Add( dp, b2Sub( b2RotateVector( dqB, rB ), b2RotateVector( dqA, rA ) ) );
		float s = c.f[kcBaseSeparation1 + o][i] + b2Dot( ds, normal );

		float velocityBias = 0.0f;
		float massScale = 1.0f;
		float impulseScale = 0.0f;
		if ( s > 0.0f )
		{
			velocityBias = s * inv_h;
		}
		else if ( useBias )
		{
			velocityBias = b2MaxFloat( softness.massScale * softness.biasRate * s, -contactSpeed );
			massScale = softness.massScale;
			impulseScale = softness.impulseScale;
		}

		b2Vec2 vrA = b2Add( vA, b2CrossSV( wA, rA ) );
		b2Vec2 vrB = b2Add( vB, b2CrossSV( wB, rB ) );
		float vn = b2Dot( b2Sub( vrB, vrA ), normal );

		float normalImpulse = c.f[kcNormalImpulse1 + o][i];
		float impulse = -c.f[kcNormalMass1 + o][i] * ( massScale * vn + velocityBias ) - impulseScale * normalImpulse;
		float newImpulse = b2MaxFloat( normalImpulse + impulse, 0.0f );
		impulse = newImpulse - normalImpulse;
		c.f[kcNormalImpulse1 + o][i] = newImpulse;
		c.f[kcTotalNormalImpulse1 + o][i] = c.f[kcTotalNormalImpulse1 + o][i] + impulse;
		totalNormalImpulse += newImpulse;

		b2Vec2 P = b2MulSV( impulse, normal );
		vA = b2MulSub( vA, mA, P );
		wA -= iA * b2Cross( rA, P );
		vB = b2MulAdd( vB, mB, P );
		wB += iB * b2Cross( rB, P );
	}

	if ( useBias == false )
	{
		float tangentSpeed = c.f[kcTangentSpeed][i];
		for ( int j = 0; j < pointCount; ++j )
		{
			int o = j == 0 ? 0 : ( kcAnchorA2X - kcAnchorA1X );
			b2Vec2 rA = { c.f[kcAnchorA1X + o][i], c.f[kcAnchorA1Y + o][i] };
			b2Vec2 rB = { c.f[kcAnchorB1X + o][i], c.f[kcAnchorB1Y + o][i] };

			b2Vec2 vrB = b2Add( vB, b2CrossSV( wB, rB ) );
			b2Vec2 vrA = b2Add( vA, b2CrossSV( wA, rA ) );
			float vt = b2Dot( b2Sub( vrB, vrA ), tangent ) - tangentSpeed;
			float impulse = c.f[kcTangentMass1 + o][i] * ( -vt );
			float maxFriction = friction * c.f[kcNormalImpulse1 + o][i];
			float tangentImpulse = c.f[kcTangentImpulse1 + o][i];
			float newImpulse = b2ClampFloat( tangentImpulse + impulse, -maxFriction, maxFriction );
			impulse = newImpulse - tangentImpulse;
			c.f[kcTangentImpulse1 + o][i] = newImpulse;

			b2Vec2 P = b2MulSV( impulse, tangent );
			vA = b2MulSub( vA, mA, P );
			wA -= iA * b2Cross( rA, P );
			vB = b2MulAdd( vB, mB, P );
			wB += iB * b2Cross( rB, P );
		}

		{
			float deltaLambda = -c.f[kcRollingMass][i] * ( wB - wA );
			float lambda = c.f[kcRollingImpulse][i];
			float maxLambda = c.f[kcRollingResistance][i] * totalNormalImpulse;
			float rolling = b2ClampFloat( lambda + deltaLambda, -maxLambda, maxLambda );
			c.f[kcRollingImpulse][i] = rolling;
			deltaLambda = rolling - lambda;
			wA -= iA * deltaLambda;
			wB += iB * deltaLambda;
		}
	}

	if ( stateA->flags & kDynamicFlag )
	{
		stateA->linearVelocity = vA;
		stateA->angularVelocity = wA;
	}
	if ( stateB->flags & kDynamicFlag )
	{
		stateB->linearVelocity = vB;
		stateB->angularVelocity = wB;
	}

	B.Store( ia, ToLane( sA ) );
	B.Store( ib, ToLane( sB ) );
}

template <typename Bodies>
B2G_DEVI void OverflowRestitution( const ConstraintSoA& c, int i, const Bodies& B, float threshold )
{
	float restitution = c.f[kcRestitution][i];
	if ( restitution == 0.0f )
	{
		return;
	}

	float mA = c.f[kcInvMassA][i], mB = c.f[kcInvMassB][i];
	float iA = c.f[kcInvIA][i], iB = c.f[kcInvIB][i];
	int ia = c.bodyA[i];
	int ib = c.bodyB[i];
	BodyState sA = ToState( B.Load( ia ) );
	BodyState* stateA = &sA;
	BodyState sB = ToState( B.Load( ib ) );
	BodyState* stateB = &sB;
	b2Vec2 vA = stateA->linearVelocity;
	float wA = stateA->angularVelocity;
	b2Vec2 vB = stateB->linearVelocity;
	float wB = stateB->angularVelocity;
	b2Vec2 normal = { c.f[kcNormalX][i], c.f[kcNormalY][i] };
	int pointCount = c.pointCount[i];

	for ( int j = 0; j < pointCount; ++j )
	{
		int o = j == 0 ? 0 : ( kcAnchorA2X - kcAnchorA1X );
		float relativeVelocity = c.f[kcRelativeVelocity1 + o][i];
		if ( relativeVelocity > -threshold || c.f[kcTotalNormalImpulse1 + o][i] == 0.0f )
		{
			continue;
		}

		b2Vec2 rA = { c.f[kcAnchorA1X + o][i], c.f[kcAnchorA1Y + o][i] };
		b2Vec2 rB = { c.f[kcAnchorB1X + o][i], c.f[kcAnchorB1Y + o][i] };
		b2Vec2 vrB = b2Add( vB, b2CrossSV( wB, rB ) );
		b2Vec2 vrA = b2Add( vA, b2CrossSV( wA, rA ) );
		float vn = b2Dot( b2Sub( vrB, vrA ), normal );

		float normalImpulse = c.f[kcNormalImpulse1 + o][i];
		float impulse = -c.f[kcNormalMass1 + o][i] * ( vn + restitution * relativeVelocity );
		float newImpulse = b2MaxFloat( normalImpulse + impulse, 0.0f );
		impulse = newImpulse - normalImpulse;
		c.f[kcNormalImpulse1 + o][i] = newImpulse;
		c.f[kcTotalNormalImpulse1 + o][i] = c.f[kcTotalNormalImpulse1 + o][i] + impulse;

		b2Vec2 P = b2MulSV( impulse, normal );
		vA = b2MulSub( vA, mA, P );
		wA -= iA * b2Cross( rA, P );
		vB = b2MulAdd( vB, mB, P );
		wB += iB * b2Cross( rB, P );
	}

	if ( stateA->flags & kDynamicFlag )
	{
		stateA->linearVelocity = vA;
		stateA->angularVelocity = wA;
	}
	if ( stateB->flags & kDynamicFlag )
	{
		stateB->linearVelocity = vB;
		stateB->angularVelocity = wB;
	}

	B.Store( ia, ToLane( sA ) );
	B.Store( ib, ToLane( sB ) );
}

} // namespace b2g

#include "joints_device.h"

namespace b2g
{

// ---------------------------------------------------------------------------------------------
// Warm start by body

// The warm-start stage of one dynamic body: the velocity deltas of its constraints' accumulated
// impulses applied in the stage's order (overflow joints, overflow contacts, then per color its
// joints and contacts). The deltas depend on impulses, anchors, normals, masses and the body's
// delta rotation, never on velocities, so the per-body sequence of float operations is exactly the
// one the color-by-color stage applies to this body (WarmStartContact, OverflowWarmStart,
// WarmStartJoint). Entries: index << 3 | owner << 2 | joint << 1 | side; the owner entry of a
// contact also accumulates the contact's total normal impulses.
template <typename Bodies>
B2G_DEVI void WarmStartJointSide( const JointSim* joint, bool sideB, const Bodies& B, b2Rot dq, b2Vec2 dp, float& vx, float& vy,
								  float& w )
{
	b2Vec2 dpOther;
	b2Rot dqOther;
	B.LoadDelta( sideB ? Bodies::JointA( joint ) : Bodies::JointB( joint ), dpOther, dqOther );
	b2Vec2 P;
	float LA, LB;
	bool active = sideB ? JointWarmStartTerms( joint, dqOther, dpOther, dq, dp, P, LA, LB )
						: JointWarmStartTerms( joint, dq, dp, dqOther, dpOther, P, LA, LB );
	if ( active == false )
	{
		return;
	}
	if ( sideB == false )
	{
		float mA = joint->invMassA, iA = joint->invIA;
		vx = vx - mA * P.x;
		vy = vy - mA * P.y;
		w -= iA * LA;
	}
	else
	{
		float mB = joint->invMassB, iB = joint->invIB;
		vx = vx + mB * P.x;
		vy = vy + mB * P.y;
		w += iB * LB;
	}
}

// kOwnerTotals: the owner entries add the contacts' normal impulses to their total normal
// impulses (otherwise each contact's own solve lane does it first, see AddWarmTotals).
// kCoherent: the impulses, written by other workgroups inside the same launch, are read with
// agent-scope loads that skip the CU's L1.
template <int kChunk = 4, bool kOwnerTotals = true, bool kCoherent = false, typename Bodies>
B2G_DEVI void WarmStartBody( const ConstraintSoA& c, const JointSim* joints, const int* jointList, int overflowContactStart,
							 const int* list, int begin, int end, const Bodies& B, b2Rot dq, b2Vec2 dp, float& vx, float& vy,
							 float& w )
{
	auto impulse = []( const float* p ) {
		return kCoherent ? __hip_atomic_load( p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT ) : *p;
	};
	// Entries go in chunks of kChunk: all of a chunk's constraint fields are gathered first
	// (independent loads), then applied in order.
	for ( int base = begin; base < end; base += kChunk )
	{
		int count = end - base < kChunk ? end - base : kChunk;
		int ent[kChunk];
#pragma unroll
		for ( int q = 0; q < kChunk; ++q )
		{
			ent[q] = q < count ? list[base + q] : 2; // padding reads as a joint entry and is skipped
		}

		float nX[kChunk], nY[kChunk], m[kChunk], iv[kChunk], roll[kChunk];
		float ni[kChunk][2], ti[kChunk][2], rx[kChunk][2], ry[kChunk][2];
		int points[kChunk];
#pragma unroll
		for ( int q = 0; q < kChunk; ++q )
		{
			if ( ent[q] & 2 )
			{
				continue;
			}
			int k = ent[q] >> 3;
			bool sideB = ( ent[q] & 1 ) != 0;
			bool overflow = k >= overflowContactStart;
			nX[q] = c.f[kcNormalX][k];
			nY[q] = c.f[kcNormalY][k];
			m[q] = sideB ? c.f[kcInvMassB][k] : c.f[kcInvMassA][k];
			iv[q] = sideB ? c.f[kcInvIB][k] : c.f[kcInvIA][k];
			roll[q] = impulse( c.f[kcRollingImpulse] + k );
			points[q] = overflow ? c.pointCount[k] : 2;
#pragma unroll
			for ( int j = 0; j < 2; ++j )
			{
				int o = j == 0 ? 0 : ( kcAnchorA2X - kcAnchorA1X );
				ni[q][j] = impulse( c.f[kcNormalImpulse1 + o] + k );
				ti[q][j] = impulse( c.f[kcTangentImpulse1 + o] + k );
				rx[q][j] = sideB ? c.f[kcAnchorB1X + o][k] : c.f[kcAnchorA1X + o][k];
				ry[q][j] = sideB ? c.f[kcAnchorB1Y + o][k] : c.f[kcAnchorA1Y + o][k];
			}
		}

#pragma unroll
		for ( int q = 0; q < kChunk; ++q )
		{
			if ( q >= count )
			{
				break;
			}
			int entry = ent[q];
			bool sideB = ( entry & 1 ) != 0;
			if ( entry & 2 )
			{
				WarmStartJointSide( joints + jointList[entry >> 3], sideB, B, dq, dp, vx, vy, w );
				continue;
			}
			bool overflow = ( entry >> 3 ) >= overflowContactStart;
			// the wide lanes form the tangent as 0 - nX, the scalar overflow solver as -nX
			float tX = nY[q];
			float tY = overflow ? -nX[q] : 0.0f - nX[q];
#pragma unroll
			for ( int j = 0; j < 2; ++j )
			{
				if ( j >= points[q] )
				{
					break;
				}
				float Px = ni[q][j] * nX[q] + ti[q][j] * tX;
				float Py = ni[q][j] * nY[q] + ti[q][j] * tY;
				if ( sideB == false )
				{
					w = w - iv[q] * ( rx[q][j] * Py - ry[q][j] * Px );
					vx = vx - m[q] * Px;
					vy = vy - m[q] * Py;
				}
				else
				{
					w = w + iv[q] * ( rx[q][j] * Py - ry[q][j] * Px );
					vx = vx + m[q] * Px;
					vy = vy + m[q] * Py;
				}
			}
			w = sideB ? w + iv[q] * roll[q] : w - iv[q] * roll[q];
		}
	}

	// total normal impulses, once per contact (kept out of the loop above so its loads carry no
	// dependence on stores)
	for ( int e = begin; kOwnerTotals && e < end; ++e )
	{
		int entry = list[e];
		if ( ( entry & 6 ) != 4 )
		{
			continue;
		}
		int k = entry >> 3;
		int points = k >= overflowContactStart ? c.pointCount[k] : 2;
		for ( int j = 0; j < points; ++j )
		{
			int o = j == 0 ? 0 : ( kcAnchorA2X - kcAnchorA1X );
			c.f[kcTotalNormalImpulse1 + o][k] = c.f[kcTotalNormalImpulse1 + o][k] + c.f[kcNormalImpulse1 + o][k];
		}
	}
}

// The warm start's contribution to a colored contact's total normal impulses, for warm starts
// that leave it to the contact's own lane (WarmStartBody with kOwnerTotals false): nothing reads
// the totals between the warm start and the contact's solve, so adding them first in the solve
// lane gives the same floats.
B2G_DEVI void AddWarmTotals( const ConstraintSoA& c, int i )
{
	c.f[kcTotalNormalImpulse1][i] = c.f[kcTotalNormalImpulse1][i] + c.f[kcNormalImpulse1][i];
	c.f[kcTotalNormalImpulse2][i] = c.f[kcTotalNormalImpulse2][i] + c.f[kcNormalImpulse2][i];
}

} // namespace b2g
