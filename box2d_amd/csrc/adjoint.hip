// Adjoint of recorded steps.
//
// One workgroup per world walks the world's records from the last step to the first. For each
// step it first recomputes the step's substep loop from the record (the checkpoint) with the same
// device functions the forward solver uses, writing the inputs of every constraint operation to a
// scratch tape, and then runs the hand-written reverse of each stage in reverse order:
//
//   finalize -> [relax, integrate positions, solve, warm start, integrate velocities] x substeps
//   -> warm-start impulse matching -> constraint preparation -> narrowphase manifold
//
// The discrete pattern of the forward (contact set, manifold features, recycled manifolds,
// warm-start matches, speculative/soft bias branches, normal and friction clamps) is held fixed:
// the reverse differentiates the branch the forward took. Reference forward code this reverses:
// body integration src/solver.c:66-162, finalize src/solver.c:553-600, warm start and solve
// src/contact_solver.c:1811-2116 (wide lanes), preparation src/contact_solver.c:1573-1809,
// manifold update and recycling src/physics_world.c:507-577 and src/contact.c:519-698, polygon
// manifold src/manifold.c:702-1028.
//
// Adjoint state carried from step to step: per body the gradients of center, rotation and
// velocities; per contact the gradients of the stored impulses (consumed by the previous step's
// solve) and of a recycled manifold's cached anchors, normal, base separations and rotations
// (consumed by the step that last recomputed the manifold).

#include <hip/hip_runtime.h>

#include "box2d_amd.h"
#include "device.h"
#include "manifold_dual.h"
#include "solver_device.h"
#include "world.h"

#include <algorithm>
#include <cstdio>
#include <cstring>

#define HIP_CHECK( expr )                                                                                                  \
	do                                                                                                                     \
	{                                                                                                                      \
		hipError_t _e = ( expr );                                                                                          \
		if ( _e != hipSuccess )                                                                                            \
		{                                                                                                                  \
			fprintf( stderr, "box2d_amd: HIP error %s at %s:%d\n", hipGetErrorString( _e ), __FILE__, __LINE__ );          \
			abort();                                                                                                       \
		}                                                                                                                  \
	}                                                                                                                      \
	while ( 0 )

namespace b2g
{

// ---------------------------------------------------------------------------------------------
// Tape storage

TapeStore::~TapeStore()
{
	for ( char* c : chunks )
	{
		hipFree( c );
	}
}

char* TapeStore::Alloc( size_t bytes )
{
	std::lock_guard<std::mutex> lock( mutex );
	bytes = ( bytes + 255 ) & ~(size_t)255;
	if ( chunks.empty() || chunkUsed + bytes > chunkSize )
	{
		size_t size = std::max<size_t>( (size_t)64 << 20, bytes );
		char* c = nullptr;
		HIP_CHECK( hipMalloc( &c, size ) );
		chunks.push_back( c );
		chunkSize = size;
		chunkUsed = 0;
	}
	char* p = chunks.back() + chunkUsed;
	chunkUsed += bytes;
	return p;
}

// ---------------------------------------------------------------------------------------------
// Scratch layouts

// inputs of one constraint operation: both bodies' lane state, then the accumulated impulses
enum OpTapeField
{
	kOpVAx, kOpVAy, kOpWA, kOpDpAx, kOpDpAy, kOpDqAc, kOpDqAs,
	kOpVBx, kOpVBy, kOpWB, kOpDpBx, kOpDpBy, kOpDqBc, kOpDqBs,
	kOpN1, kOpN2, kOpT1, kOpT2,
	kOpFieldCount
};

// per-constraint adjoints of the step's constraint quantities
enum ConstraintAdjField
{
	kaRA1X, kaRA1Y, kaRB1X, kaRB1Y,
	kaRA2X, kaRA2Y, kaRB2X, kaRB2Y,
	kaNX, kaNY,
	kaBase1, kaBase2,
	kaNormalMass1, kaNormalMass2,
	kaTangentMass1, kaTangentMass2,
	kaFriction,
	kaN1, kaN2, kaT1, kaT2,
	kaSep1, kaSep2,
	kaFieldCount
};

// per-contact (by contact id) accumulators carried between steps
enum ContactAccField
{
	kAccN1, kAccN2, kAccT1, kAccT2, // stored-impulse gradients
	kAccImpulseTag,					// step whose stored impulses they belong to (int bits)
	kAccA1X, kAccA1Y, kAccA2X, kAccA2Y,
	kAccB1X, kAccB1Y, kAccB2X, kAccB2Y,
	kAccNX, kAccNY,
	kAccBase1, kAccBase2,
	kAccQAc, kAccQAs, kAccQBc, kAccQBs,
	kAccCacheValid, // recycled-manifold cache gradients pending (int bits)
	kAccFieldCount
};

struct AdjointWorldArgs
{
	const char* const* steps; // stepCount record addresses (device)
	const char* final;
	int stepCount;
	int nbMax;
	int ncMax;
	int contactCap;
	int subStepMax;
	const ShapeDev* shapes;
	const BodySim* sims;

	float* opTape;	   // subStepMax * 3 * kOpFieldCount * ncMax, then subStepMax * 3 * nbMax
	float* cfields;	   // kcFieldCount * ncMax
	int* cints;		   // 6 * ncMax
	float* cadj;	   // kaFieldCount * ncMax
	float* contactAcc; // kAccFieldCount * contactCap

	const b2amd_BodyState* finalGrad; // by body id
	b2amd_BodyState* initialGrad;	  // by body id
	float* frictionGrad;			  // by shape id
	int* diag;						  // [0] replay mismatches, [1] unsupported bits
};

constexpr int kAdjointThreads = 256;

B2G_DEVI const b2amd_TapeStep* StepHeader( const char* r )
{
	return reinterpret_cast<const b2amd_TapeStep*>( r );
}

B2G_DEVI const b2amd_TapeBody* StepBodies( const char* r )
{
	return reinterpret_cast<const b2amd_TapeBody*>( r + kTapeBodiesOffset );
}

B2G_DEVI const b2amd_TapeConstraint* StepConstraints( const char* r, int bodyCount )
{
	size_t off = kTapeBodiesOffset + ( ( (size_t)bodyCount * sizeof( b2amd_TapeBody ) + 255 ) & ~(size_t)255 );
	return reinterpret_cast<const b2amd_TapeConstraint*>( r + off );
}

B2G_DEVI bool SameBits( float a, float b )
{
	return __float_as_uint( a ) == __float_as_uint( b );
}

// Adjoint body state in LDS, by slot (slot 0: static bodies, never written).
struct AdjBodies
{
	float* vx;
	float* vy;
	float* w;
	float* dpx;
	float* dpy;
	float* dqc;
	float* dqs;
	float* cx;
	float* cy;
	float* qc;
	float* qs;
};

// d( R(q) v )/dq contracted with the adjoint g of the rotated vector
B2G_DEVI void RotGrad( float gx, float gy, b2Vec2 v, float* qc, float* qs )
{
	*qc += gx * v.x + gy * v.y;
	*qs += -gx * v.y + gy * v.x;
}

// ---------------------------------------------------------------------------------------------
// Reverse of the constraint operations (colored wide-lane arithmetic of solver_device.h)

struct LaneAdj
{
	float vAx, vAy, wA, vBx, vBy, wB;
};

struct GeomAdj
{
	float rA[2][2], rB[2][2];
	float nX, nY, tX, tY;
	float base[2], nm[2], tm[2];
	float friction;
	float dpX, dpY;
	float dqAc, dqAs, dqBc, dqBs;
};

struct ConstraintConsts
{
	float mA, mB, iA, iB;
	float nX, nY, tX, tY;
	float rA[2][2], rB[2][2];
	float base[2], nm[2], tm[2];
	float friction, tangentSpeed;
	float biasRate, massScale, impulseScale;
	int pointCount;
};

B2G_DEVI ConstraintConsts LoadConsts( const ConstraintSoA& c, int i, bool useBias )
{
	ConstraintConsts k;
	k.mA = c.f[kcInvMassA][i];
	k.mB = c.f[kcInvMassB][i];
	k.iA = c.f[kcInvIA][i];
	k.iB = c.f[kcInvIB][i];
	k.nX = c.f[kcNormalX][i];
	k.nY = c.f[kcNormalY][i];
	k.tX = k.nY;
	k.tY = 0.0f - k.nX;
	for ( int j = 0; j < 2; ++j )
	{
		int o = j == 0 ? 0 : ( kcAnchorA2X - kcAnchorA1X );
		k.rA[j][0] = c.f[kcAnchorA1X + o][i];
		k.rA[j][1] = c.f[kcAnchorA1Y + o][i];
		k.rB[j][0] = c.f[kcAnchorB1X + o][i];
		k.rB[j][1] = c.f[kcAnchorB1Y + o][i];
		k.base[j] = c.f[kcBaseSeparation1 + o][i];
		k.nm[j] = c.f[kcNormalMass1 + o][i];
		k.tm[j] = c.f[kcTangentMass1 + o][i];
	}
	k.friction = c.f[kcFriction][i];
	k.tangentSpeed = c.f[kcTangentSpeed][i];
	if ( useBias )
	{
		k.biasRate = c.f[kcMassScale][i] * c.f[kcBiasRate][i];
		k.massScale = c.f[kcMassScale][i];
		k.impulseScale = c.f[kcImpulseScale][i];
	}
	else
	{
		k.biasRate = 0.0f;
		k.massScale = 1.0f;
		k.impulseScale = 0.0f;
	}
	k.pointCount = c.pointCount[i];
	return k;
}

struct LaneVel
{
	float vAx, vAy, wA, vBx, vBy, wB;
};

// Velocity change of an impulse P along (px, py): the update every contact operation applies.
B2G_DEVI void ApplyImpulse( LaneVel& v, const ConstraintConsts& k, int j, float Px, float Py )
{
	const float* rA = k.rA[j];
	const float* rB = k.rB[j];
	v.vAx = v.vAx - k.mA * Px;
	v.vAy = v.vAy - k.mA * Py;
	v.wA = v.wA - k.iA * ( rA[0] * Py - rA[1] * Px );
	v.vBx = v.vBx + k.mB * Px;
	v.vBy = v.vBy + k.mB * Py;
	v.wB = v.wB + k.iB * ( rB[0] * Py - rB[1] * Px );
}

// Reverse of ApplyImpulse: returns the impulse adjoint, accumulates the anchor adjoints.
B2G_DEVI void ApplyImpulseRev( const LaneAdj& a, const ConstraintConsts& k, int j, float Px, float Py, GeomAdj& g, float* Pbx,
							   float* Pby )
{
	const float* rA = k.rA[j];
	const float* rB = k.rB[j];
	float gA = -k.iA * a.wA;
	float gB = k.iB * a.wB;
	*Pbx = -k.mA * a.vAx + k.mB * a.vBx + gA * ( -rA[1] ) + gB * ( -rB[1] );
	*Pby = -k.mA * a.vAy + k.mB * a.vBy + gA * rA[0] + gB * rB[0];
	g.rA[j][0] += gA * Py;
	g.rA[j][1] += gA * ( -Px );
	g.rB[j][0] += gB * Py;
	g.rB[j][1] += gB * ( -Px );
}

// Reverse of the relative velocity dv = (vB + wB x rB) - (vA + wA x rA) at the pre-op velocities.
B2G_DEVI void RelVelRev( LaneAdj& a, const LaneVel& v, const ConstraintConsts& k, int j, float dvxb, float dvyb, GeomAdj& g )
{
	const float* rA = k.rA[j];
	const float* rB = k.rB[j];
	a.vBx += dvxb;
	a.vAx -= dvxb;
	a.vBy += dvyb;
	a.vAy -= dvyb;
	a.wB += -rB[1] * dvxb + rB[0] * dvyb;
	a.wA += rA[1] * dvxb - rA[0] * dvyb;
	g.rB[j][1] += -v.wB * dvxb;
	g.rB[j][0] += v.wB * dvyb;
	g.rA[j][1] += v.wA * dvxb;
	g.rA[j][0] += -v.wA * dvyb;
}

struct NormalFwd
{
	float s, dsx, dsy, vn, dvx, dvy, bias, pms, pis, lambda, delta;
	bool mask, softActive, active;
};

// Forward of one normal point (SolveNormalPoint), velocities updated in place.
B2G_DEVI NormalFwd NormalPointFwd( LaneVel& v, const ConstraintConsts& k, int j, float lambda, float dpX, float dpY, float dqAc,
								   float dqAs, float dqBc, float dqBs, float inv_h, float contactSpeed )
{
	NormalFwd f;
	const float* rA = k.rA[j];
	const float* rB = k.rB[j];
	float rsAx = dqAc * rA[0] - dqAs * rA[1];
	float rsAy = dqAs * rA[0] + dqAc * rA[1];
	float rsBx = dqBc * rB[0] - dqBs * rB[1];
	float rsBy = dqBs * rB[0] + dqBc * rB[1];
	f.dsx = dpX + ( rsBx - rsAx );
	f.dsy = dpY + ( rsBy - rsAy );
	f.s = ( k.nX * f.dsx + k.nY * f.dsy ) + k.base[j];
	f.mask = f.s > 0.0f;
	float specBias = f.s * inv_h;
	float soft = k.biasRate * f.s;
	f.softActive = soft >= contactSpeed;
	float softBias = f.softActive ? soft : contactSpeed;
	f.bias = f.mask ? specBias : softBias;
	f.pms = f.mask ? 1.0f : k.massScale;
	f.pis = f.mask ? 0.0f : k.impulseScale;
	f.dvx = ( v.vBx - v.wB * rB[1] ) - ( v.vAx - v.wA * rA[1] );
	f.dvy = ( v.vBy + v.wB * rB[0] ) - ( v.vAy + v.wA * rA[0] );
	f.vn = f.dvx * k.nX + f.dvy * k.nY;
	float neg = k.nm[j] * ( f.pms * f.vn + f.bias ) + f.pis * lambda;
	float cand = lambda - neg;
	f.active = cand >= 0.0f;
	float newImpulse = f.active ? cand : 0.0f;
	f.lambda = newImpulse;
	f.delta = newImpulse - lambda;
	ApplyImpulse( v, k, j, f.delta * k.nX, f.delta * k.nY );
	return f;
}

// Reverse of one normal point. a: adjoint of the post-point velocities (updated to pre-point);
// lb: adjoint of the accumulated impulse after the point (replaced by the one before it).
B2G_DEVI void NormalPointRev( LaneAdj& a, const LaneVel& pre, const NormalFwd& f, const ConstraintConsts& k, int j, float dqAc,
							  float dqAs, float dqBc, float dqBs, float inv_h, float* lb, GeomAdj& g )
{
	float Px = f.delta * k.nX;
	float Py = f.delta * k.nY;
	float Pbx, Pby;
	ApplyImpulseRev( a, k, j, Px, Py, g, &Pbx, &Pby );
	float deltab = Pbx * k.nX + Pby * k.nY;
	g.nX += f.delta * Pbx;
	g.nY += f.delta * Pby;
	float newb = *lb + deltab;
	float oldb = -deltab;
	float candb = f.active ? newb : 0.0f;
	oldb += candb;
	float negb = -candb;
	g.nm[j] += negb * ( f.pms * f.vn + f.bias );
	float vnb = negb * k.nm[j] * f.pms;
	float biasb = negb * k.nm[j];
	oldb += negb * f.pis;
	float sb = f.mask ? biasb * inv_h : ( f.softActive ? biasb * k.biasRate : 0.0f );

	float dvxb = vnb * k.nX;
	float dvyb = vnb * k.nY;
	g.nX += vnb * f.dvx;
	g.nY += vnb * f.dvy;
	RelVelRev( a, pre, k, j, dvxb, dvyb, g );

	g.nX += sb * f.dsx;
	g.nY += sb * f.dsy;
	g.base[j] += sb;
	float dsxb = sb * k.nX;
	float dsyb = sb * k.nY;
	g.dpX += dsxb;
	g.dpY += dsyb;
	const float* rA = k.rA[j];
	const float* rB = k.rB[j];
	// rsB = R(dqB) rB with adjoint ds; rsA = R(dqA) rA with adjoint -ds
	g.dqBc += dsxb * rB[0] + dsyb * rB[1];
	g.dqBs += -dsxb * rB[1] + dsyb * rB[0];
	g.rB[j][0] += dsxb * dqBc + dsyb * dqBs;
	g.rB[j][1] += -dsxb * dqBs + dsyb * dqBc;
	g.dqAc -= dsxb * rA[0] + dsyb * rA[1];
	g.dqAs -= -dsxb * rA[1] + dsyb * rA[0];
	g.rA[j][0] -= dsxb * dqAc + dsyb * dqAs;
	g.rA[j][1] -= -dsxb * dqAs + dsyb * dqAc;
	*lb = oldb;
}

struct FrictionFwd
{
	float vt, dvx, dvy, maxF, delta;
	bool innerIsCand, resIsLower;
};

B2G_DEVI FrictionFwd FrictionPointFwd( LaneVel& v, const ConstraintConsts& k, int j, float lambdaT, float lambdaN )
{
	FrictionFwd f;
	const float* rA = k.rA[j];
	const float* rB = k.rB[j];
	f.dvx = ( v.vBx - v.wB * rB[1] ) - ( v.vAx - v.wA * rA[1] );
	f.dvy = ( v.vBy + v.wB * rB[0] ) - ( v.vAy + v.wA * rA[0] );
	float vt = f.dvx * k.tX + f.dvy * k.tY;
	f.vt = vt - k.tangentSpeed;
	float neg = k.tm[j] * f.vt;
	f.maxF = k.friction * lambdaN;
	float cand = lambdaT - neg;
	f.innerIsCand = cand <= f.maxF;
	float inner = f.innerIsCand ? cand : f.maxF;
	float lower = 0.0f - f.maxF;
	f.resIsLower = lower >= inner;
	float newImpulse = f.resIsLower ? lower : inner;
	f.delta = newImpulse - lambdaT;
	ApplyImpulse( v, k, j, f.delta * k.tX, f.delta * k.tY );
	return f;
}

// Reverse of one friction point; lbT: adjoint of the tangent impulse (after -> before), lbN:
// adjoint of this point's normal impulse (receives the friction-bound term), lambdaN its value.
B2G_DEVI void FrictionPointRev( LaneAdj& a, const LaneVel& pre, const FrictionFwd& f, const ConstraintConsts& k, int j, float lambdaN,
								float* lbT, float* lbN, GeomAdj& g )
{
	float Px = f.delta * k.tX;
	float Py = f.delta * k.tY;
	float Pbx, Pby;
	ApplyImpulseRev( a, k, j, Px, Py, g, &Pbx, &Pby );
	float deltab = Pbx * k.tX + Pby * k.tY;
	g.tX += f.delta * Pbx;
	g.tY += f.delta * Pby;
	float newb = *lbT + deltab;
	float oldb = -deltab;
	float candb = 0.0f, maxFb = 0.0f;
	if ( f.resIsLower )
	{
		maxFb = -newb;
	}
	else if ( f.innerIsCand )
	{
		candb = newb;
	}
	else
	{
		maxFb = newb;
	}
	oldb += candb;
	float negb = -candb;
	g.tm[j] += negb * f.vt;
	float vtb = negb * k.tm[j];
	float dvxb = vtb * k.tX;
	float dvyb = vtb * k.tY;
	g.tX += vtb * f.dvx;
	g.tY += vtb * f.dvy;
	RelVelRev( a, pre, k, j, dvxb, dvyb, g );
	g.friction += maxFb * lambdaN;
	*lbN += maxFb * k.friction;
	*lbT = oldb;
}

// Reverse of SolveContact (useBias: solve stage, else relax with friction) for constraint i.
B2G_DEVI void SolveContactRev( const ConstraintSoA& c, int i, const float* tp, int stride, bool useBias, float inv_h,
							   float contactSpeed, const AdjBodies& AB, FieldTable adj )
{
	ConstraintConsts k = LoadConsts( c, i, useBias );
	int pc = k.pointCount;
	int sa = c.bodyA[i];
	int sb = c.bodyB[i];

	LaneVel v;
	v.vAx = tp[kOpVAx * stride];
	v.vAy = tp[kOpVAy * stride];
	v.wA = tp[kOpWA * stride];
	v.vBx = tp[kOpVBx * stride];
	v.vBy = tp[kOpVBy * stride];
	v.wB = tp[kOpWB * stride];
	float dpX = tp[kOpDpBx * stride] - tp[kOpDpAx * stride];
	float dpY = tp[kOpDpBy * stride] - tp[kOpDpAy * stride];
	float dqAc = tp[kOpDqAc * stride], dqAs = tp[kOpDqAs * stride];
	float dqBc = tp[kOpDqBc * stride], dqBs = tp[kOpDqBs * stride];
	float ln[2] = { tp[kOpN1 * stride], tp[kOpN2 * stride] };
	float lt[2] = { tp[kOpT1 * stride], tp[kOpT2 * stride] };

	// forward replay, keeping the pre-velocities of every sub-operation
	LaneVel preN[2], preT[2];
	NormalFwd nf[2];
	FrictionFwd ff[2];
	float lnNew[2] = { ln[0], ln[1] };
	for ( int j = 0; j < pc; ++j )
	{
		preN[j] = v;
		nf[j] = NormalPointFwd( v, k, j, ln[j], dpX, dpY, dqAc, dqAs, dqBc, dqBs, inv_h, contactSpeed );
		lnNew[j] = nf[j].lambda;
	}
	if ( useBias == false )
	{
		for ( int j = 0; j < pc; ++j )
		{
			preT[j] = v;
			ff[j] = FrictionPointFwd( v, k, j, lt[j], lnNew[j] );
		}
	}

	// reverse
	LaneAdj a;
	a.vAx = sa != kStaticSlot ? AB.vx[sa] : 0.0f;
	a.vAy = sa != kStaticSlot ? AB.vy[sa] : 0.0f;
	a.wA = sa != kStaticSlot ? AB.w[sa] : 0.0f;
	a.vBx = sb != kStaticSlot ? AB.vx[sb] : 0.0f;
	a.vBy = sb != kStaticSlot ? AB.vy[sb] : 0.0f;
	a.wB = sb != kStaticSlot ? AB.w[sb] : 0.0f;

	GeomAdj g;
	memset( &g, 0, sizeof( g ) );
	float lbN[2] = { adj[kaN1][i], adj[kaN2][i] };
	float lbT[2] = { adj[kaT1][i], adj[kaT2][i] };

	if ( useBias == false )
	{
		for ( int j = pc - 1; j >= 0; --j )
		{
			FrictionPointRev( a, preT[j], ff[j], k, j, lnNew[j], &lbT[j], &lbN[j], g );
		}
	}
	for ( int j = pc - 1; j >= 0; --j )
	{
		NormalPointRev( a, preN[j], nf[j], k, j, dqAc, dqAs, dqBc, dqBs, inv_h, &lbN[j], g );
	}

	// tangent = (n.y, -n.x)
	g.nY += g.tX;
	g.nX -= g.tY;

	adj[kaN1][i] = lbN[0];
	adj[kaN2][i] = lbN[1];
	adj[kaT1][i] = lbT[0];
	adj[kaT2][i] = lbT[1];
	adj[kaRA1X][i] += g.rA[0][0];
	adj[kaRA1Y][i] += g.rA[0][1];
	adj[kaRB1X][i] += g.rB[0][0];
	adj[kaRB1Y][i] += g.rB[0][1];
	adj[kaRA2X][i] += g.rA[1][0];
	adj[kaRA2Y][i] += g.rA[1][1];
	adj[kaRB2X][i] += g.rB[1][0];
	adj[kaRB2Y][i] += g.rB[1][1];
	adj[kaNX][i] += g.nX;
	adj[kaNY][i] += g.nY;
	adj[kaBase1][i] += g.base[0];
	adj[kaBase2][i] += g.base[1];
	adj[kaNormalMass1][i] += g.nm[0];
	adj[kaNormalMass2][i] += g.nm[1];
	adj[kaTangentMass1][i] += g.tm[0];
	adj[kaTangentMass2][i] += g.tm[1];
	adj[kaFriction][i] += g.friction;

	if ( sa != kStaticSlot )
	{
		AB.vx[sa] = a.vAx;
		AB.vy[sa] = a.vAy;
		AB.w[sa] = a.wA;
		AB.dpx[sa] -= g.dpX;
		AB.dpy[sa] -= g.dpY;
		AB.dqc[sa] += g.dqAc;
		AB.dqs[sa] += g.dqAs;
	}
	if ( sb != kStaticSlot )
	{
		AB.vx[sb] = a.vBx;
		AB.vy[sb] = a.vBy;
		AB.w[sb] = a.wB;
		AB.dpx[sb] += g.dpX;
		AB.dpy[sb] += g.dpY;
		AB.dqc[sb] += g.dqBc;
		AB.dqs[sb] += g.dqBs;
	}
}

// Reverse of WarmStartContact (linear in the impulses and anchors; velocities pass through).
B2G_DEVI void WarmStartContactRev( const ConstraintSoA& c, int i, const float* tp, int stride, const AdjBodies& AB, FieldTable adj )
{
	ConstraintConsts k = LoadConsts( c, i, true );
	int pc = k.pointCount;
	int sa = c.bodyA[i];
	int sb = c.bodyB[i];
	LaneAdj a;
	a.vAx = sa != kStaticSlot ? AB.vx[sa] : 0.0f;
	a.vAy = sa != kStaticSlot ? AB.vy[sa] : 0.0f;
	a.wA = sa != kStaticSlot ? AB.w[sa] : 0.0f;
	a.vBx = sb != kStaticSlot ? AB.vx[sb] : 0.0f;
	a.vBy = sb != kStaticSlot ? AB.vy[sb] : 0.0f;
	a.wB = sb != kStaticSlot ? AB.w[sb] : 0.0f;
	float ln[2] = { tp[kOpN1 * stride], tp[kOpN2 * stride] };
	float lt[2] = { tp[kOpT1 * stride], tp[kOpT2 * stride] };

	GeomAdj g;
	memset( &g, 0, sizeof( g ) );
	float lbN[2] = { 0.0f, 0.0f }, lbT[2] = { 0.0f, 0.0f };
	for ( int j = 0; j < pc; ++j )
	{
		float Px = ln[j] * k.nX + lt[j] * k.tX;
		float Py = ln[j] * k.nY + lt[j] * k.tY;
		float Pbx, Pby;
		ApplyImpulseRev( a, k, j, Px, Py, g, &Pbx, &Pby );
		lbN[j] = Pbx * k.nX + Pby * k.nY;
		lbT[j] = Pbx * k.tX + Pby * k.tY;
		g.nX += ln[j] * Pbx;
		g.nY += ln[j] * Pby;
		g.tX += lt[j] * Pbx;
		g.tY += lt[j] * Pby;
	}
	g.nY += g.tX;
	g.nX -= g.tY;

	adj[kaN1][i] += lbN[0];
	adj[kaN2][i] += lbN[1];
	adj[kaT1][i] += lbT[0];
	adj[kaT2][i] += lbT[1];
	adj[kaRA1X][i] += g.rA[0][0];
	adj[kaRA1Y][i] += g.rA[0][1];
	adj[kaRB1X][i] += g.rB[0][0];
	adj[kaRB1Y][i] += g.rB[0][1];
	adj[kaRA2X][i] += g.rA[1][0];
	adj[kaRA2Y][i] += g.rA[1][1];
	adj[kaRB2X][i] += g.rB[1][0];
	adj[kaRB2Y][i] += g.rB[1][1];
	adj[kaNX][i] += g.nX;
	adj[kaNY][i] += g.nY;
}

// ---------------------------------------------------------------------------------------------
// Reverse of constraint preparation and of the manifold update

// Body pose of a constraint side: awake bodies from the step's record, static bodies from the
// world's body array (they do not move).
struct Pose
{
	b2Vec2 center;
	b2Vec2 origin;
	b2Rot q;
	b2Vec2 localCenter;
};

B2G_DEVI Pose PoseOf( int slot, int bodyId, const b2amd_TapeBody* bodies, const BodySim* sims )
{
	Pose p;
	if ( slot != kStaticSlot )
	{
		const b2amd_TapeBody& b = bodies[slot - 1];
		p.center = b.center;
		p.origin = b.origin;
		p.q = b.rotation;
		p.localCenter = b.localCenter;
	}
	else
	{
		const BodySim& s = sims[bodyId];
		p.center = s.center;
		p.origin = s.transform.p;
		p.q = s.transform.q;
		p.localCenter = s.localCenter;
	}
	return p;
}

// Pose gradients of one side: center, origin (transform position) and rotation.
struct PoseAdj
{
	float cx, cy, px, py, qc, qs;
};

// Reverse of the full manifold update of a polygon pair (CollideBody's b2UpdateContact path).
// Inputs: adjoints of the world normal, the center-relative anchors and the separations. Returns
// false when the recomputed values differ from the record.
__device__ __noinline__ bool FullManifoldRev( const b2amd_TapeConstraint& r, const ShapeDev& shapeA, const ShapeDev& shapeB,
											  const Pose& A, const Pose& B, const float nb[2], const float rAb[2][2],
											  const float rBb[2][2], const float sepb[2], PoseAdj* gA, PoseAdj* gB )
{
	// relative transform (b2InvMulWorldTransforms)
	b2Rot qA = A.q, qB = B.q;
	b2Rot rq;
	rq.s = qA.c * qB.s - qA.s * qB.c;
	rq.c = qA.c * qB.c + qA.s * qB.s;
	b2Vec2 d = { B.origin.x - A.origin.x, B.origin.y - A.origin.y };
	b2Vec2 rp = { qA.c * d.x + qA.s * d.y, -qA.s * d.x + qA.c * d.y };

	DV2 xfp = { D4{ rp.x, { 1.0f, 0.0f, 0.0f, 0.0f } }, D4{ rp.y, { 0.0f, 1.0f, 0.0f, 0.0f } } };
	DV2 xfq = { D4{ rq.c, { 0.0f, 0.0f, 1.0f, 0.0f } }, D4{ rq.s, { 0.0f, 0.0f, 0.0f, 1.0f } } };
	DManifold m = DCollidePolygons( shapeA.polygon, shapeB.polygon, xfp, xfq );

	// world values and the check against the record
	bool same = m.pointCount == r.pointCount;
	b2Vec2 ln = { m.normal.x.v, m.normal.y.v };
	b2Vec2 normal = b2RotateVector( qA, ln );
	same = same && SameBits( normal.x, r.normal.x ) && SameBits( normal.y, r.normal.y );
	b2Vec2 originDelta = { A.origin.x - B.origin.x, A.origin.y - B.origin.y };
	b2Vec2 coA = b2RotateVector( qA, A.localCenter );
	b2Vec2 coB = b2RotateVector( qB, B.localCenter );
	int pc = same ? m.pointCount : 0;
	for ( int j = 0; j < pc; ++j )
	{
		b2Vec2 lp = { m.points[j].point.x.v, m.points[j].point.y.v };
		b2Vec2 aw = b2RotateVector( qA, lp );
		b2Vec2 bw = b2Add( aw, originDelta );
		b2Vec2 anchorA = b2Sub( aw, coA );
		b2Vec2 anchorB = b2Sub( bw, coB );
		same = same && m.points[j].id == r.id[j] && SameBits( anchorA.x, r.anchorA[j].x ) && SameBits( anchorA.y, r.anchorA[j].y ) &&
			   SameBits( anchorB.x, r.anchorB[j].x ) && SameBits( anchorB.y, r.anchorB[j].y ) &&
			   SameBits( m.points[j].separation.v, r.separation[j] );
	}

	// reverse of the world conversion
	float lnb[2] = { qA.c * nb[0] + qA.s * nb[1], -qA.s * nb[0] + qA.c * nb[1] };
	RotGrad( nb[0], nb[1], ln, &gA->qc, &gA->qs );
	float relb[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
	for ( int k = 0; k < 4; ++k )
	{
		relb[k] += lnb[0] * m.normal.x.d[k] + lnb[1] * m.normal.y.d[k];
	}
	for ( int j = 0; j < r.pointCount && j < m.pointCount; ++j )
	{
		float awb[2] = { rAb[j][0] + rBb[j][0], rAb[j][1] + rBb[j][1] };
		// originDelta = pA - pB
		gA->px += rBb[j][0];
		gA->py += rBb[j][1];
		gB->px -= rBb[j][0];
		gB->py -= rBb[j][1];
		// centerOffsets
		RotGrad( -rAb[j][0], -rAb[j][1], A.localCenter, &gA->qc, &gA->qs );
		RotGrad( -rBb[j][0], -rBb[j][1], B.localCenter, &gB->qc, &gB->qs );
		// aw = R(qA) lp
		b2Vec2 lp = { m.points[j].point.x.v, m.points[j].point.y.v };
		RotGrad( awb[0], awb[1], lp, &gA->qc, &gA->qs );
		float lpb[2] = { qA.c * awb[0] + qA.s * awb[1], -qA.s * awb[0] + qA.c * awb[1] };
		for ( int k = 0; k < 4; ++k )
		{
			relb[k] += lpb[0] * m.points[j].point.x.d[k] + lpb[1] * m.points[j].point.y.d[k] + sepb[j] * m.points[j].separation.d[k];
		}
	}

	// relative position rp = R(qA)^T (pB - pA)
	float rpb[2] = { relb[0], relb[1] };
	float db[2] = { qA.c * rpb[0] - qA.s * rpb[1], qA.s * rpb[0] + qA.c * rpb[1] };
	gA->qc += rpb[0] * d.x + rpb[1] * d.y;
	gA->qs += rpb[0] * d.y - rpb[1] * d.x;
	gB->px += db[0];
	gB->py += db[1];
	gA->px -= db[0];
	gA->py -= db[1];
	// relative rotation
	float rqcb = relb[2], rqsb = relb[3];
	gA->qc += rqsb * qB.s + rqcb * qB.c;
	gA->qs += -rqsb * qB.c + rqcb * qB.s;
	gB->qs += rqsb * qA.c + rqcb * qA.s;
	gB->qc += -rqsb * qA.s + rqcb * qA.c;
	return same;
}

// Reverse of the recycled update: separation = base + dot(dc + R(dqB) aB - R(dqA) aA, normal),
// dq = q * inverse(cached q), dc = centerB - centerA. Cache gradients are written to cacheb
// (kAccA1X.. layout offset by kAccA1X). Returns false when the separation differs from the record.
B2G_DEVI bool RecycledManifoldRev( const b2amd_TapeConstraint& r, const Pose& A, const Pose& B, const float sepb[2], float* cacheb,
								   PoseAdj* gA, PoseAdj* gB )
{
	b2Rot cqA = r.cachedRotationA, cqB = r.cachedRotationB;
	b2Rot dqA = b2MulRot( A.q, b2InvertRot( cqA ) );
	b2Rot dqB = b2MulRot( B.q, b2InvertRot( cqB ) );
	b2Vec2 normal = r.normal;
	b2Vec2 dc = b2SubPos( B.center, A.center );
	bool same = true;
	float dqAb[2] = { 0.0f, 0.0f }, dqBb[2] = { 0.0f, 0.0f };
	for ( int j = 0; j < r.pointCount; ++j )
	{
		b2Vec2 aA = r.anchorA[j], aB = r.anchorB[j];
		b2Vec2 rA = b2RotateVector( dqA, aA );
		b2Vec2 rB = b2RotateVector( dqB, aB );
		b2Vec2 dp = b2Add( dc, b2Sub( rB, rA ) );
		float sep = r.baseSeparation[j] + b2Dot( dp, normal );
		same = same && SameBits( sep, r.separation[j] );

		float s = sepb[j];
		cacheb[kAccBase1 - kAccA1X + j] += s;
		cacheb[kAccNX - kAccA1X] += s * dp.x;
		cacheb[kAccNY - kAccA1X] += s * dp.y;
		float dpb[2] = { s * normal.x, s * normal.y };
		gB->cx += dpb[0];
		gB->cy += dpb[1];
		gA->cx -= dpb[0];
		gA->cy -= dpb[1];
		// rB = R(dqB) aB with adjoint dpb; rA with -dpb
		dqBb[0] += dpb[0] * aB.x + dpb[1] * aB.y;
		dqBb[1] += -dpb[0] * aB.y + dpb[1] * aB.x;
		cacheb[kAccB1X - kAccA1X + 2 * j] += dpb[0] * dqB.c + dpb[1] * dqB.s;
		cacheb[kAccB1Y - kAccA1X + 2 * j] += -dpb[0] * dqB.s + dpb[1] * dqB.c;
		dqAb[0] -= dpb[0] * aA.x + dpb[1] * aA.y;
		dqAb[1] -= -dpb[0] * aA.y + dpb[1] * aA.x;
		cacheb[kAccA1X - kAccA1X + 2 * j] -= dpb[0] * dqA.c + dpb[1] * dqA.s;
		cacheb[kAccA1Y - kAccA1X + 2 * j] -= -dpb[0] * dqA.s + dpb[1] * dqA.c;
	}
	// dq = (q.c cq.c + q.s cq.s, q.s cq.c - q.c cq.s) as (c, s)
	gA->qc += dqAb[0] * cqA.c - dqAb[1] * cqA.s;
	gA->qs += dqAb[0] * cqA.s + dqAb[1] * cqA.c;
	cacheb[kAccQAc - kAccA1X] += dqAb[0] * A.q.c + dqAb[1] * A.q.s;
	cacheb[kAccQAs - kAccA1X] += dqAb[0] * A.q.s - dqAb[1] * A.q.c;
	gB->qc += dqBb[0] * cqB.c - dqBb[1] * cqB.s;
	gB->qs += dqBb[0] * cqB.s + dqBb[1] * cqB.c;
	cacheb[kAccQBc - kAccA1X] += dqBb[0] * B.q.c + dqBb[1] * B.q.s;
	cacheb[kAccQBs - kAccA1X] += dqBb[0] * B.q.s - dqBb[1] * B.q.c;
	return same;
}

// ---------------------------------------------------------------------------------------------
// The backward kernel

__global__ __launch_bounds__( kAdjointThreads ) void AdjointKernel( const AdjointWorldArgs* worlds )
{
	const AdjointWorldArgs& W = worlds[blockIdx.x];
	extern __shared__ char s_lds[];
	__shared__ int s_unsupported;
	__shared__ int s_mismatch;

	const int cap = W.nbMax + 1;
	float* f = reinterpret_cast<float*>( s_lds );
	LdsBodies B{ f, f + cap, f + 2 * cap, f + 3 * cap, f + 4 * cap, f + 5 * cap, f + 6 * cap, nullptr };
	AdjBodies AB{ f + 7 * cap, f + 8 * cap, f + 9 * cap, f + 10 * cap, f + 11 * cap, f + 12 * cap,
				  f + 13 * cap, f + 14 * cap, f + 15 * cap, f + 16 * cap, f + 17 * cap };
	B.flags = reinterpret_cast<uint16_t*>( f + 18 * cap );

	const int ncMax = W.ncMax;
	ConstraintSoA c;
	c.f = FieldTable{ W.cfields, ncMax };
	c.bodyA = W.cints;
	c.bodyB = W.cints + ncMax;
	c.contactId = W.cints + 2 * ncMax;
	c.pointCount = W.cints + 3 * ncMax;
	c.slotA = W.cints + 4 * ncMax;
	c.slotB = W.cints + 5 * ncMax;
	FieldTable adj{ W.cadj, ncMax };
	FieldTable acc{ W.contactAcc, W.contactCap };

	if ( threadIdx.x == 0 )
	{
		s_unsupported = 0;
		s_mismatch = 0;
	}
	for ( int id = threadIdx.x; id < W.contactCap; id += blockDim.x )
	{
		acc[kAccImpulseTag][id] = __int_as_float( -1 );
		acc[kAccCacheValid][id] = __int_as_float( 0 );
	}

	// final-state gradient into the adjoint slots
	{
		const b2amd_TapeBody* fb = StepBodies( W.final );
		int nb = StepHeader( W.final )->bodyCount;
		for ( int s = threadIdx.x; s < cap; s += blockDim.x )
		{
			float g[7] = { 0, 0, 0, 0, 0, 0, 0 };
			if ( s >= 1 && s <= nb )
			{
				const b2amd_BodyState& fg = W.finalGrad[fb[s - 1].bodyId];
				g[0] = fg.center.x;
				g[1] = fg.center.y;
				g[2] = fg.rotation.c;
				g[3] = fg.rotation.s;
				g[4] = fg.linearVelocity.x;
				g[5] = fg.linearVelocity.y;
				g[6] = fg.angularVelocity;
			}
			AB.cx[s] = g[0];
			AB.cy[s] = g[1];
			AB.qc[s] = g[2];
			AB.qs[s] = g[3];
			AB.vx[s] = g[4];
			AB.vy[s] = g[5];
			AB.w[s] = g[6];
			AB.dpx[s] = 0.0f;
			AB.dpy[s] = 0.0f;
			AB.dqc[s] = 0.0f;
			AB.dqs[s] = 0.0f;
		}
	}
	__syncthreads();

	for ( int k = W.stepCount - 1; k >= 0; --k )
	{
		const char* rec = W.steps[k];
		const char* next = k + 1 < W.stepCount ? W.steps[k + 1] : W.final;
		const b2amd_TapeStep hdr = *StepHeader( rec );
		const int nb = hdr.bodyCount;
		const int nc = hdr.constraintCount;
		const int S = hdr.subStepCount;
		const b2amd_TapeBody* bodies = StepBodies( rec );
		const b2amd_TapeConstraint* cons = StepConstraints( rec, nb );
		const b2amd_TapeBody* nextBodies = StepBodies( next );

		StepParams p;
		memset( &p, 0, sizeof( p ) );
		p.dt = hdr.timeStep;
		p.inv_dt = hdr.inv_dt;
		p.h = hdr.h;
		p.inv_h = hdr.inv_h;
		p.maxLinearVelocity = hdr.maxLinearSpeed;
		p.gravity = hdr.gravity;
		const float contactSpeed = -hdr.contactSpeed;
		const float warmStartScale = hdr.enableWarmStarting ? 1.0f : 0.0f;

		if ( threadIdx.x == 0 )
		{
			int bits = 0;
			bits |= hdr.overflowCount > 0 ? 1 : 0;
			bits |= hdr.jointCount > 0 ? 2 : 0;
			bits |= StepHeader( next )->bodyCount != nb ? 32 : 0;
			atomicOr( &s_unsupported, bits );
		}

		// ---- forward recompute of the step from its record ----
		for ( int s = threadIdx.x; s < cap; s += blockDim.x )
		{
			if ( s == 0 )
			{
				B.vx[0] = 0.0f;
				B.vy[0] = 0.0f;
				B.w[0] = 0.0f;
				B.flags[0] = 0;
			}
			else if ( s <= nb )
			{
				const b2amd_TapeBody& b = bodies[s - 1];
				B.vx[s] = b.linearVelocity.x;
				B.vy[s] = b.linearVelocity.y;
				B.w[s] = b.angularVelocity;
				B.flags[s] = (uint16_t)b.flags;
				int bits = 0;
				bits |= ( b.flags & kAllLocks ) ? 64 : 0;
				bits |= ( b.flags & ( kIsFast | kHadTimeOfImpact | kIsBullet ) ) ? 128 : 0;
				bits |= nextBodies[s - 1].bodyId != b.bodyId ? 32 : 0;
				if ( bits )
				{
					atomicOr( &s_unsupported, bits );
				}
			}
			B.dpx[s] = 0.0f;
			B.dpy[s] = 0.0f;
			B.dqc[s] = 1.0f;
			B.dqs[s] = 0.0f;
		}

		// constraint preparation from the record (PrepareContactsBody arithmetic)
		for ( int i = threadIdx.x; i < nc; i += blockDim.x )
		{
			const b2amd_TapeConstraint& r = cons[i];
			c.bodyA[i] = r.slotA;
			c.bodyB[i] = r.slotB;
			c.slotA[i] = r.slotA;
			c.slotB[i] = r.slotB;
			c.contactId[i] = r.contactId;
			c.pointCount[i] = r.pointCount;
			float mA = r.slotA != kStaticSlot ? bodies[r.slotA - 1].invMass : 0.0f;
			float iA = r.slotA != kStaticSlot ? bodies[r.slotA - 1].invInertia : 0.0f;
			float mB = r.slotB != kStaticSlot ? bodies[r.slotB - 1].invMass : 0.0f;
			float iB = r.slotB != kStaticSlot ? bodies[r.slotB - 1].invInertia : 0.0f;
			c.f[kcInvMassA][i] = mA;
			c.f[kcInvMassB][i] = mB;
			c.f[kcInvIA][i] = iA;
			c.f[kcInvIB][i] = iB;
			{
				float kk = iA + iB;
				c.f[kcRollingMass][i] = kk > 0.0f ? 1.0f / kk : 0.0f;
			}
			b2Vec2 normal = r.normal;
			c.f[kcNormalX][i] = normal.x;
			c.f[kcNormalY][i] = normal.y;
			c.f[kcFriction][i] = r.friction;
			c.f[kcTangentSpeed][i] = r.tangentSpeed;
			c.f[kcRestitution][i] = r.restitution;
			c.f[kcRollingResistance][i] = r.rollingResistance;
			c.f[kcRollingImpulse][i] = 0.0f;
			c.f[kcBiasRate][i] = r.biasRate;
			c.f[kcMassScale][i] = r.massScale;
			c.f[kcImpulseScale][i] = r.impulseScale;
			b2Vec2 tangent = b2RightPerp( normal );
			for ( int j = 0; j < 2; ++j )
			{
				int o = j == 0 ? 0 : ( kcAnchorA2X - kcAnchorA1X );
				if ( j < r.pointCount )
				{
					b2Vec2 rA = r.anchorA[j];
					b2Vec2 rB = r.anchorB[j];
					c.f[kcAnchorA1X + o][i] = rA.x;
					c.f[kcAnchorA1Y + o][i] = rA.y;
					c.f[kcAnchorB1X + o][i] = rB.x;
					c.f[kcAnchorB1Y + o][i] = rB.y;
					c.f[kcBaseSeparation1 + o][i] = r.separation[j] - b2Dot( b2Sub( rB, rA ), normal );
					c.f[kcNormalImpulse1 + o][i] = r.normalImpulse[j];
					c.f[kcTangentImpulse1 + o][i] = r.tangentImpulse[j];
					c.f[kcTotalNormalImpulse1 + o][i] = 0.0f;
					float rnA = b2Cross( rA, normal );
					float rnB = b2Cross( rB, normal );
					float kNormal = mA + mB + iA * rnA * rnA + iB * rnB * rnB;
					c.f[kcNormalMass1 + o][i] = kNormal > 0.0f ? 1.0f / kNormal : 0.0f;
					float rtA = b2Cross( rA, tangent );
					float rtB = b2Cross( rB, tangent );
					float kTangent = mA + mB + iA * rtA * rtA + iB * rtB * rtB;
					c.f[kcTangentMass1 + o][i] = kTangent > 0.0f ? 1.0f / kTangent : 0.0f;
					c.f[kcRelativeVelocity1 + o][i] = 0.0f;
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
			int bits = 0;
			bits |= ( r.restitution != 0.0f || r.rollingResistance != 0.0f ) ? 4 : 0;
			bits |= ( W.shapes[r.shapeIdA].type != b2_polygonShape || W.shapes[r.shapeIdB].type != b2_polygonShape ) ? 8 : 0;
			if ( bits )
			{
				atomicOr( &s_unsupported, bits );
			}
		}
		__syncthreads();

		const int colors = hdr.activeColorCount;
		auto opTape = [&]( int sub, int stage ) { return W.opTape + (size_t)( ( sub * 3 + stage ) * kOpFieldCount ) * ncMax; };
		float* posTape = W.opTape + (size_t)W.subStepMax * 3 * kOpFieldCount * ncMax;
		auto record = [&]( float* t, int i ) {
			int sa = c.bodyA[i], sb = c.bodyB[i];
			t[kOpVAx * ncMax + i] = B.vx[sa];
			t[kOpVAy * ncMax + i] = B.vy[sa];
			t[kOpWA * ncMax + i] = B.w[sa];
			t[kOpDpAx * ncMax + i] = B.dpx[sa];
			t[kOpDpAy * ncMax + i] = B.dpy[sa];
			t[kOpDqAc * ncMax + i] = B.dqc[sa];
			t[kOpDqAs * ncMax + i] = B.dqs[sa];
			t[kOpVBx * ncMax + i] = B.vx[sb];
			t[kOpVBy * ncMax + i] = B.vy[sb];
			t[kOpWB * ncMax + i] = B.w[sb];
			t[kOpDpBx * ncMax + i] = B.dpx[sb];
			t[kOpDpBy * ncMax + i] = B.dpy[sb];
			t[kOpDqBc * ncMax + i] = B.dqc[sb];
			t[kOpDqBs * ncMax + i] = B.dqs[sb];
			t[kOpN1 * ncMax + i] = c.f[kcNormalImpulse1][i];
			t[kOpN2 * ncMax + i] = c.f[kcNormalImpulse2][i];
			t[kOpT1 * ncMax + i] = c.f[kcTangentImpulse1][i];
			t[kOpT2 * ncMax + i] = c.f[kcTangentImpulse2][i];
		};

		for ( int sub = 0; sub < S; ++sub )
		{
			for ( int s = threadIdx.x; s < nb; s += blockDim.x )
			{
				const b2amd_TapeBody& b = bodies[s];
				BodySim sim;
				sim.invMass = b.invMass;
				sim.invInertia = b.invInertia;
				sim.force = b.force;
				sim.torque = b.torque;
				sim.linearDamping = b.linearDamping;
				sim.angularDamping = b.angularDamping;
				sim.gravityScale = b.gravityScale;
				BodyState st = ToState( B.Load( s + 1 ) );
				IntegrateVelocity( sim, &st, p );
				B.vx[s + 1] = st.linearVelocity.x;
				B.vy[s + 1] = st.linearVelocity.y;
				B.w[s + 1] = st.angularVelocity;
			}
			__syncthreads();

			for ( int stage = 0; stage < 2; ++stage )
			{
				float* t = opTape( sub, stage );
				for ( int color = 0; color < colors; ++color )
				{
					int start = hdr.colorStart[color];
					int count = hdr.colorStart[color + 1] - start;
					for ( int q = threadIdx.x; q < count; q += blockDim.x )
					{
						int i = start + q;
						record( t, i );
						if ( stage == 0 )
						{
							WarmStartContact( c, i, B );
						}
						else
						{
							SolveContact( c, i, B, true, p.inv_h, contactSpeed, false );
						}
					}
					__syncthreads();
				}
			}

			for ( int s = threadIdx.x; s < nb; s += blockDim.x )
			{
				float* pt = posTape + (size_t)sub * 3 * W.nbMax;
				pt[s] = B.w[s + 1];
				pt[W.nbMax + s] = B.dqc[s + 1];
				pt[2 * W.nbMax + s] = B.dqs[s + 1];
				BodyState st = ToState( B.Load( s + 1 ) );
				IntegratePosition( &st, p );
				if ( st.flags & kIsSpeedCapped )
				{
					atomicOr( &s_unsupported, 16 );
				}
				B.vx[s + 1] = st.linearVelocity.x;
				B.vy[s + 1] = st.linearVelocity.y;
				B.w[s + 1] = st.angularVelocity;
				B.dpx[s + 1] = st.deltaPosition.x;
				B.dpy[s + 1] = st.deltaPosition.y;
				B.dqc[s + 1] = st.deltaRotation.c;
				B.dqs[s + 1] = st.deltaRotation.s;
			}
			__syncthreads();

			{
				float* t = opTape( sub, 2 );
				for ( int color = 0; color < colors; ++color )
				{
					int start = hdr.colorStart[color];
					int count = hdr.colorStart[color + 1] - start;
					for ( int q = threadIdx.x; q < count; q += blockDim.x )
					{
						int i = start + q;
						record( t, i );
						SolveContact( c, i, B, false, p.inv_h, contactSpeed, false );
					}
					__syncthreads();
				}
			}
		}

		// finalize check against the next record, then the reverse of finalize
		for ( int s = threadIdx.x; s < nb; s += blockDim.x )
		{
			const b2amd_TapeBody& b = bodies[s];
			b2Rot dq = { B.dqc[s + 1], B.dqs[s + 1] };
			b2Vec2 dp = { B.dpx[s + 1], B.dpy[s + 1] };
			b2Vec2 center = b2OffsetPos( b.center, dp );
			b2Rot m = b2MulRot( dq, b.rotation );
			b2Rot q1 = b2NormalizeRot( m );
			const b2amd_TapeBody& nbd = nextBodies[s];
			bool same = SameBits( center.x, nbd.center.x ) && SameBits( center.y, nbd.center.y ) &&
						SameBits( q1.c, nbd.rotation.c ) && SameBits( q1.s, nbd.rotation.s ) &&
						SameBits( B.vx[s + 1], nbd.linearVelocity.x ) && SameBits( B.vy[s + 1], nbd.linearVelocity.y ) &&
						SameBits( B.w[s + 1], nbd.angularVelocity );
			if ( same == false )
			{
				atomicAdd( &s_mismatch, 1 );
			}

			int slot = s + 1;
			float cb0 = AB.cx[slot], cb1 = AB.cy[slot];
			float qb0 = AB.qc[slot], qb1 = AB.qs[slot];
			// q1 = m / |m|
			float mag = sqrtf( m.s * m.s + m.c * m.c );
			float inv = mag > 0.0f ? 1.0f / mag : 0.0f;
			float yy = q1.c * qb0 + q1.s * qb1;
			float mbc = ( qb0 - q1.c * yy ) * inv;
			float mbs = ( qb1 - q1.s * yy ) * inv;
			b2Rot q = b.rotation;
			AB.dqc[slot] = mbs * q.s + mbc * q.c;
			AB.dqs[slot] = mbs * q.c - mbc * q.s;
			AB.qc[slot] = mbs * dq.s + mbc * dq.c;
			AB.qs[slot] = mbs * dq.c - mbc * dq.s;
			AB.dpx[slot] = cb0;
			AB.dpy[slot] = cb1;
			// center passes through (cx, cy unchanged); velocities pass through
		}

		// stored impulses: their gradients were left by the next step's warm start
		for ( int i = threadIdx.x; i < nc; i += blockDim.x )
		{
			int cid = cons[i].contactId;
			bool mine = __float_as_int( acc[kAccImpulseTag][cid] ) == k;
			adj[kaN1][i] = mine ? acc[kAccN1][cid] : 0.0f;
			adj[kaN2][i] = mine ? acc[kAccN2][cid] : 0.0f;
			adj[kaT1][i] = mine ? acc[kAccT1][cid] : 0.0f;
			adj[kaT2][i] = mine ? acc[kAccT2][cid] : 0.0f;
			for ( int fld = 0; fld < kaFieldCount; ++fld )
			{
				if ( fld < kaN1 || fld > kaT2 )
				{
					adj[fld][i] = 0.0f;
				}
			}
		}
		__syncthreads();

		// ---- reverse of the substeps ----
		for ( int sub = S - 1; sub >= 0; --sub )
		{
			{
				const float* t = opTape( sub, 2 );
				for ( int color = colors - 1; color >= 0; --color )
				{
					int start = hdr.colorStart[color];
					int count = hdr.colorStart[color + 1] - start;
					for ( int q = threadIdx.x; q < count; q += blockDim.x )
					{
						int i = start + q;
						SolveContactRev( c, i, t + i, ncMax, false, p.inv_h, contactSpeed, AB, adj );
					}
					__syncthreads();
				}
			}

			for ( int s = threadIdx.x; s < nb; s += blockDim.x )
			{
				int slot = s + 1;
				const float* pt = posTape + (size_t)sub * 3 * W.nbMax;
				float w = pt[s];
				float dqc = pt[W.nbMax + s];
				float dqs = pt[2 * W.nbMax + s];
				float h = p.h;
				AB.vx[slot] += h * AB.dpx[slot];
				AB.vy[slot] += h * AB.dpy[slot];
				float a = h * w;
				float q2c = dqc - a * dqs;
				float q2s = dqs + a * dqc;
				float mag = sqrtf( q2s * q2s + q2c * q2c );
				float inv = mag > 0.0f ? 1.0f / mag : 0.0f;
				float yc = q2c * inv, ys = q2s * inv;
				float ybc = AB.dqc[slot], ybs = AB.dqs[slot];
				float yy = yc * ybc + ys * ybs;
				float q2bc = ( ybc - yc * yy ) * inv;
				float q2bs = ( ybs - ys * yy ) * inv;
				AB.dqc[slot] = q2bc + a * q2bs;
				AB.dqs[slot] = -a * q2bc + q2bs;
				float ab = -dqs * q2bc + dqc * q2bs;
				AB.w[slot] += h * ab;
			}
			__syncthreads();

			for ( int stage = 1; stage >= 0; --stage )
			{
				const float* t = opTape( sub, stage );
				for ( int color = colors - 1; color >= 0; --color )
				{
					int start = hdr.colorStart[color];
					int count = hdr.colorStart[color + 1] - start;
					for ( int q = threadIdx.x; q < count; q += blockDim.x )
					{
						int i = start + q;
						if ( stage == 1 )
						{
							SolveContactRev( c, i, t + i, ncMax, true, p.inv_h, contactSpeed, AB, adj );
						}
						else
						{
							WarmStartContactRev( c, i, t + i, ncMax, AB, adj );
						}
					}
					__syncthreads();
				}
			}

			for ( int s = threadIdx.x; s < nb; s += blockDim.x )
			{
				const b2amd_TapeBody& b = bodies[s];
				int slot = s + 1;
				float linearDamping = 1.0f / ( 1.0f + p.h * b.linearDamping );
				float angularDamping = 1.0f / ( 1.0f + p.h * b.angularDamping );
				AB.vx[slot] *= linearDamping;
				AB.vy[slot] *= linearDamping;
				AB.w[slot] *= angularDamping;
			}
			__syncthreads();
		}

		// ---- warm-start sources, preparation and manifold ----
		for ( int i = threadIdx.x; i < nc; i += blockDim.x )
		{
			const b2amd_TapeConstraint& r = cons[i];
			int cid = r.contactId;
			float lb[2][2] = { { adj[kaN1][i], adj[kaT1][i] }, { adj[kaN2][i], adj[kaT2][i] } };
			float out[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
			bool any = false;
			for ( int j = 0; j < r.pointCount; ++j )
			{
				int src = ( r.info >> ( 2 + 2 * j ) ) & 3;
				if ( src > 0 )
				{
					out[src - 1] += warmStartScale * lb[j][0];
					out[2 + src - 1] += warmStartScale * lb[j][1];
					any = true;
				}
			}
			acc[kAccN1][cid] = out[0];
			acc[kAccN2][cid] = out[1];
			acc[kAccT1][cid] = out[2];
			acc[kAccT2][cid] = out[3];
			acc[kAccImpulseTag][cid] = __int_as_float( any ? k - 1 : -1 );

			// reverse of the preparation: anchors, normal and separations of the manifold
			float nX = r.normal.x, nY = r.normal.y;
			float tX = nY, tY = 0.0f - nX;
			float iA = c.f[kcInvIA][i], iB = c.f[kcInvIB][i];
			float nb_[2] = { adj[kaNX][i], adj[kaNY][i] };
			float tb[2] = { 0.0f, 0.0f };
			float rAb[2][2] = { { adj[kaRA1X][i], adj[kaRA1Y][i] }, { adj[kaRA2X][i], adj[kaRA2Y][i] } };
			float rBb[2][2] = { { adj[kaRB1X][i], adj[kaRB1Y][i] }, { adj[kaRB2X][i], adj[kaRB2Y][i] } };
			float sepb[2] = { 0.0f, 0.0f };
			for ( int j = 0; j < r.pointCount; ++j )
			{
				int o = j == 0 ? 0 : ( kcAnchorA2X - kcAnchorA1X );
				b2Vec2 rA = r.anchorA[j], rB = r.anchorB[j];
				float bb = adj[kaBase1 + j][i];
				sepb[j] += bb;
				float dx = rB.x - rA.x, dy = rB.y - rA.y;
				nb_[0] -= bb * dx;
				nb_[1] -= bb * dy;
				rBb[j][0] -= bb * nX;
				rBb[j][1] -= bb * nY;
				rAb[j][0] += bb * nX;
				rAb[j][1] += bb * nY;

				float nm = c.f[kcNormalMass1 + o][i];
				if ( nm > 0.0f )
				{
					float kb = -adj[kaNormalMass1 + j][i] * nm * nm;
					float rnA = rA.x * nY - rA.y * nX;
					float rnB = rB.x * nY - rB.y * nX;
					float rnAb = kb * 2.0f * iA * rnA;
					float rnBb = kb * 2.0f * iB * rnB;
					rAb[j][0] += rnAb * nY;
					rAb[j][1] -= rnAb * nX;
					nb_[1] += rnAb * rA.x;
					nb_[0] -= rnAb * rA.y;
					rBb[j][0] += rnBb * nY;
					rBb[j][1] -= rnBb * nX;
					nb_[1] += rnBb * rB.x;
					nb_[0] -= rnBb * rB.y;
				}
				float tm = c.f[kcTangentMass1 + o][i];
				if ( tm > 0.0f )
				{
					float kb = -adj[kaTangentMass1 + j][i] * tm * tm;
					float rtA = rA.x * tY - rA.y * tX;
					float rtB = rB.x * tY - rB.y * tX;
					float rtAb = kb * 2.0f * iA * rtA;
					float rtBb = kb * 2.0f * iB * rtB;
					rAb[j][0] += rtAb * tY;
					rAb[j][1] -= rtAb * tX;
					tb[1] += rtAb * rA.x;
					tb[0] -= rtAb * rA.y;
					rBb[j][0] += rtBb * tY;
					rBb[j][1] -= rtBb * tX;
					tb[1] += rtBb * rB.x;
					tb[0] -= rtBb * rB.y;
				}
			}
			nb_[1] += tb[0];
			nb_[0] -= tb[1];
			adj[kaNX][i] = nb_[0];
			adj[kaNY][i] = nb_[1];
			adj[kaRA1X][i] = rAb[0][0];
			adj[kaRA1Y][i] = rAb[0][1];
			adj[kaRA2X][i] = rAb[1][0];
			adj[kaRA2Y][i] = rAb[1][1];
			adj[kaRB1X][i] = rBb[0][0];
			adj[kaRB1Y][i] = rBb[0][1];
			adj[kaRB2X][i] = rBb[1][0];
			adj[kaRB2Y][i] = rBb[1][1];
			adj[kaSep1][i] = sepb[0];
			adj[kaSep2][i] = sepb[1];
		}
		__syncthreads();

		// manifold reverse, one color at a time so the two bodies of a lane are its own
		for ( int color = 0; color < colors; ++color )
		{
			int start = hdr.colorStart[color];
			int count = hdr.colorStart[color + 1] - start;
			for ( int q = threadIdx.x; q < count; q += blockDim.x )
			{
				int i = start + q;
				const b2amd_TapeConstraint& r = cons[i];
				int cid = r.contactId;
				Pose A = PoseOf( r.slotA, r.bodyIdA, bodies, W.sims );
				Pose Bp = PoseOf( r.slotB, r.bodyIdB, bodies, W.sims );
				PoseAdj gA, gB;
				memset( &gA, 0, sizeof( gA ) );
				memset( &gB, 0, sizeof( gB ) );
				float nb_[2] = { adj[kaNX][i], adj[kaNY][i] };
				float rAb[2][2] = { { adj[kaRA1X][i], adj[kaRA1Y][i] }, { adj[kaRA2X][i], adj[kaRA2Y][i] } };
				float rBb[2][2] = { { adj[kaRB1X][i], adj[kaRB1Y][i] }, { adj[kaRB2X][i], adj[kaRB2Y][i] } };
				float sepb[2] = { adj[kaSep1][i], adj[kaSep2][i] };
				bool same = true;
				if ( r.info & 1u )
				{
					// recycled: the anchors and normal are the cache's; the cache's gradients wait
					// for the step that computed it
					float cacheb[kAccCacheValid - kAccA1X];
					bool valid = __float_as_int( acc[kAccCacheValid][cid] ) != 0;
					for ( int fld = kAccA1X; fld < kAccCacheValid; ++fld )
					{
						cacheb[fld - kAccA1X] = valid ? acc[fld][cid] : 0.0f;
					}
					for ( int j = 0; j < 2; ++j )
					{
						cacheb[kAccA1X - kAccA1X + 2 * j] += rAb[j][0];
						cacheb[kAccA1Y - kAccA1X + 2 * j] += rAb[j][1];
						cacheb[kAccB1X - kAccA1X + 2 * j] += rBb[j][0];
						cacheb[kAccB1Y - kAccA1X + 2 * j] += rBb[j][1];
					}
					cacheb[kAccNX - kAccA1X] += nb_[0];
					cacheb[kAccNY - kAccA1X] += nb_[1];
					same = RecycledManifoldRev( r, A, Bp, sepb, cacheb, &gA, &gB );
					for ( int fld = kAccA1X; fld < kAccCacheValid; ++fld )
					{
						acc[fld][cid] = cacheb[fld - kAccA1X];
					}
					acc[kAccCacheValid][cid] = __int_as_float( 1 );
				}
				else
				{
					if ( __float_as_int( acc[kAccCacheValid][cid] ) != 0 )
					{
						for ( int j = 0; j < 2; ++j )
						{
							rAb[j][0] += acc[kAccA1X + 2 * j][cid];
							rAb[j][1] += acc[kAccA1Y + 2 * j][cid];
							rBb[j][0] += acc[kAccB1X + 2 * j][cid];
							rBb[j][1] += acc[kAccB1Y + 2 * j][cid];
							sepb[j] += acc[kAccBase1 + j][cid];
						}
						nb_[0] += acc[kAccNX][cid];
						nb_[1] += acc[kAccNY][cid];
						gA.qc += acc[kAccQAc][cid];
						gA.qs += acc[kAccQAs][cid];
						gB.qc += acc[kAccQBc][cid];
						gB.qs += acc[kAccQBs][cid];
						acc[kAccCacheValid][cid] = __int_as_float( 0 );
					}
					const ShapeDev& shapeA = W.shapes[r.shapeIdA];
					const ShapeDev& shapeB = W.shapes[r.shapeIdB];
					if ( shapeA.type == b2_polygonShape && shapeB.type == b2_polygonShape )
					{
						same = FullManifoldRev( r, shapeA, shapeB, A, Bp, nb_, rAb, rBb, sepb, &gA, &gB );
					}
				}
				if ( same == false )
				{
					atomicAdd( &s_mismatch, 1 );
				}

				// origin = center - R(q) localCenter
				gA.cx += gA.px;
				gA.cy += gA.py;
				RotGrad( -gA.px, -gA.py, A.localCenter, &gA.qc, &gA.qs );
				gB.cx += gB.px;
				gB.cy += gB.py;
				RotGrad( -gB.px, -gB.py, Bp.localCenter, &gB.qc, &gB.qs );
				if ( r.slotA != kStaticSlot )
				{
					AB.cx[r.slotA] += gA.cx;
					AB.cy[r.slotA] += gA.cy;
					AB.qc[r.slotA] += gA.qc;
					AB.qs[r.slotA] += gA.qs;
				}
				if ( r.slotB != kStaticSlot )
				{
					AB.cx[r.slotB] += gB.cx;
					AB.cy[r.slotB] += gB.cy;
					AB.qc[r.slotB] += gB.qc;
					AB.qs[r.slotB] += gB.qs;
				}
			}
			__syncthreads();
		}

		// contact friction sqrt(frictionA * frictionB) -> shape frictions (one lane, fixed order)
		if ( threadIdx.x == 0 )
		{
			for ( int i = 0; i < nc; ++i )
			{
				const b2amd_TapeConstraint& r = cons[i];
				float fb = adj[kaFriction][i];
				float fr = r.friction;
				if ( fb != 0.0f && fr > 0.0f )
				{
					float fA = W.shapes[r.shapeIdA].friction;
					float fB = W.shapes[r.shapeIdB].friction;
					W.frictionGrad[r.shapeIdA] += fb * 0.5f * fB / fr;
					W.frictionGrad[r.shapeIdB] += fb * 0.5f * fA / fr;
				}
			}
		}
		__syncthreads();
	}

	// gradient with respect to the initial state
	{
		const b2amd_TapeBody* b0 = StepBodies( W.steps[0] );
		int nb = StepHeader( W.steps[0] )->bodyCount;
		for ( int s = threadIdx.x; s < nb; s += blockDim.x )
		{
			int slot = s + 1;
			b2amd_BodyState g;
			g.center = { AB.cx[slot], AB.cy[slot] };
			g.rotation = { AB.qc[slot], AB.qs[slot] };
			g.linearVelocity = { AB.vx[slot], AB.vy[slot] };
			g.angularVelocity = AB.w[slot];
			g.padding = 0.0f;
			W.initialGrad[b0[s].bodyId] = g;
		}
	}
	__syncthreads();
	if ( threadIdx.x == 0 )
	{
		atomicAdd( W.diag + 0, s_mismatch );
		atomicOr( W.diag + 1, s_unsupported );
	}
}

// ---------------------------------------------------------------------------------------------
// Host side

struct TapeHost
{
	TapeStore store;
	std::vector<uint32_t> worldKeys; // index1 | generation << 16
	std::vector<const ShapeDev*> shapes;
	std::vector<const BodySim*> sims;
	std::vector<int> contactCaps;
	int bodyStride = 0;
	int shapeStride = 0;
};

static uint32_t WorldKey( b2WorldId id )
{
	return (uint32_t)id.index1 | ( (uint32_t)id.generation << 16 );
}

static World* WorldOf( uint32_t key )
{
	b2WorldId id;
	id.index1 = (uint16_t)( key & 0xFFFF );
	id.generation = (uint16_t)( key >> 16 );
	return GetWorld( id );
}

static void Detach( TapeHost* t )
{
	for ( size_t i = 0; i < t->worldKeys.size(); ++i )
	{
		World* w = WorldOf( t->worldKeys[i] );
		if ( w != nullptr && w->device->Tape() == &t->store )
		{
			w->device->SetTape( nullptr, 0 );
		}
	}
}

} // namespace b2g

using namespace b2g;

extern "C" {

b2amd_Tape* b2amd_Tape_Create( const b2WorldId* worldIds, int count )
{
	TapeHost* t = new TapeHost();
	t->store.worlds.resize( count );
	for ( int i = 0; i < count; ++i )
	{
		World* w = GetWorld( worldIds[i] );
		if ( w == nullptr || w->device->Tape() != nullptr )
		{
			Detach( t );
			delete t;
			return nullptr;
		}
		t->worldKeys.push_back( WorldKey( worldIds[i] ) );
		t->store.worlds[i].world = w;
		w->device->SetTape( &t->store, i );
	}
	return reinterpret_cast<b2amd_Tape*>( t );
}

int b2amd_Tape_Finish( b2amd_Tape* tape )
{
	TapeHost* t = reinterpret_cast<TapeHost*>( tape );
	if ( t->store.finished )
	{
		return t->store.worlds.empty() ? 0 : (int)t->store.worlds[0].steps.size();
	}
	int n = (int)t->worldKeys.size();
	t->shapes.assign( n, nullptr );
	t->sims.assign( n, nullptr );
	t->contactCaps.assign( n, 1 );
	for ( int i = 0; i < n; ++i )
	{
		World* w = WorldOf( t->worldKeys[i] );
		TapeStore::WorldTape& wt = t->store.worlds[i];
		if ( w == nullptr )
		{
			continue;
		}
		int nb = w->device->AwakeBodyCount();
		char* dst = t->store.Alloc( TapeRecordBytes( nb, 0 ) );
		w->device->CaptureFinalState( dst, nb );
		wt.final = dst;
		wt.finalBodyCount = nb;
		w->device->SetTape( nullptr, 0 );
		t->shapes[i] = static_cast<const ShapeDev*>( w->device->ShapesDevice() );
		t->sims[i] = static_cast<const BodySim*>( w->device->SimsDevice() );
		t->contactCaps[i] = std::max<int>( 1, (int)w->contacts.size() );
		t->bodyStride = std::max<int>( t->bodyStride, (int)w->bodies.size() );
		t->shapeStride = std::max<int>( t->shapeStride, (int)w->shapes.size() );
	}
	t->store.finished = true;
	return n > 0 ? (int)t->store.worlds[0].steps.size() : 0;
}

void b2amd_Tape_Destroy( b2amd_Tape* tape )
{
	TapeHost* t = reinterpret_cast<TapeHost*>( tape );
	if ( t == nullptr )
	{
		return;
	}
	if ( t->store.finished == false )
	{
		Detach( t );
	}
	delete t;
}

int b2amd_Tape_GetStepCount( b2amd_Tape* tape, int worldIndex )
{
	TapeHost* t = reinterpret_cast<TapeHost*>( tape );
	if ( worldIndex < 0 || worldIndex >= (int)t->store.worlds.size() )
	{
		return 0;
	}
	return (int)t->store.worlds[worldIndex].steps.size();
}

int b2amd_Tape_GetBodyStride( b2amd_Tape* tape )
{
	return reinterpret_cast<TapeHost*>( tape )->bodyStride;
}

int b2amd_Tape_GetShapeStride( b2amd_Tape* tape )
{
	return reinterpret_cast<TapeHost*>( tape )->shapeStride;
}

int b2amd_Tape_ReadStep( b2amd_Tape* tape, int worldIndex, int step, b2amd_TapeStep* header, b2amd_TapeBody* bodies, int bodyCapacity,
						 b2amd_TapeConstraint* constraints, int constraintCapacity )
{
	TapeHost* t = reinterpret_cast<TapeHost*>( tape );
	if ( worldIndex < 0 || worldIndex >= (int)t->store.worlds.size() )
	{
		return -1;
	}
	const TapeStore::WorldTape& wt = t->store.worlds[worldIndex];
	int steps = (int)wt.steps.size();
	const char* rec = nullptr;
	if ( step >= 0 && step < steps )
	{
		rec = wt.steps[step];
	}
	else if ( step == steps && wt.final != nullptr )
	{
		rec = wt.final;
	}
	if ( rec == nullptr )
	{
		return -1;
	}
	b2amd_TapeStep h;
	HIP_CHECK( hipMemcpy( &h, rec, sizeof( h ), hipMemcpyDeviceToHost ) );
	if ( header != nullptr )
	{
		*header = h;
	}
	int nb = std::min( h.bodyCount, bodyCapacity );
	if ( bodies != nullptr && nb > 0 )
	{
		HIP_CHECK( hipMemcpy( bodies, rec + kTapeBodiesOffset, nb * sizeof( b2amd_TapeBody ), hipMemcpyDeviceToHost ) );
	}
	int nc = std::min( h.constraintCount, constraintCapacity );
	if ( constraints != nullptr && nc > 0 )
	{
		HIP_CHECK( hipMemcpy( constraints, rec + TapeConstraintsOffset( h.bodyCount ), nc * sizeof( b2amd_TapeConstraint ),
							  hipMemcpyDeviceToHost ) );
	}
	return 0;
}

int b2amd_Tape_GetFinalStates( b2amd_Tape* tape, b2amd_BodyState* states )
{
	TapeHost* t = reinterpret_cast<TapeHost*>( tape );
	if ( t->store.finished == false )
	{
		return -1;
	}
	std::vector<b2amd_TapeBody> bodies;
	for ( size_t i = 0; i < t->store.worlds.size(); ++i )
	{
		const TapeStore::WorldTape& wt = t->store.worlds[i];
		if ( wt.final == nullptr )
		{
			continue;
		}
		bodies.resize( std::max( wt.finalBodyCount, 1 ) );
		if ( wt.finalBodyCount > 0 )
		{
			HIP_CHECK( hipMemcpy( bodies.data(), wt.final + kTapeBodiesOffset, wt.finalBodyCount * sizeof( b2amd_TapeBody ),
								  hipMemcpyDeviceToHost ) );
		}
		for ( int k = 0; k < wt.finalBodyCount; ++k )
		{
			const b2amd_TapeBody& b = bodies[k];
			b2amd_BodyState& s = states[i * (size_t)t->bodyStride + b.bodyId];
			s.center = b.center;
			s.rotation = b.rotation;
			s.linearVelocity = b.linearVelocity;
			s.angularVelocity = b.angularVelocity;
			s.padding = 0.0f;
		}
	}
	return 0;
}

int b2amd_Tape_GetBody( b2amd_Tape* tape, int worldIndex, int bodyId, b2amd_TapeBody* body )
{
	TapeHost* t = reinterpret_cast<TapeHost*>( tape );
	if ( worldIndex < 0 || worldIndex >= (int)t->worldKeys.size() )
	{
		return -1;
	}
	World* w = WorldOf( t->worldKeys[worldIndex] );
	if ( w == nullptr || bodyId < 0 || bodyId >= (int)w->bodies.size() )
	{
		return -1;
	}
	BodySim sim;
	const BodySim* sims = static_cast<const BodySim*>( w->device->SimsDevice() );
	HIP_CHECK( hipMemcpy( &sim, sims + bodyId, sizeof( sim ), hipMemcpyDeviceToHost ) );
	memset( body, 0, sizeof( *body ) );
	body->bodyId = bodyId;
	body->flags = sim.flags;
	body->center = sim.center;
	body->origin = sim.transform.p;
	body->rotation = sim.transform.q;
	body->invMass = sim.invMass;
	body->invInertia = sim.invInertia;
	body->gravityScale = sim.gravityScale;
	body->linearDamping = sim.linearDamping;
	body->angularDamping = sim.angularDamping;
	body->localCenter = sim.localCenter;
	body->force = sim.force;
	body->torque = sim.torque;
	return 0;
}

int b2amd_Tape_GetShape( b2amd_Tape* tape, int worldIndex, int shapeId, b2amd_Shape* shape, int* bodyId, float* friction )
{
	TapeHost* t = reinterpret_cast<TapeHost*>( tape );
	if ( worldIndex < 0 || worldIndex >= (int)t->worldKeys.size() )
	{
		return -1;
	}
	World* w = WorldOf( t->worldKeys[worldIndex] );
	if ( w == nullptr || shapeId < 0 || shapeId >= (int)w->shapeData.size() )
	{
		return -1;
	}
	const ShapeDev& s = w->shapeData[shapeId];
	memset( shape, 0, sizeof( *shape ) );
	shape->type = s.type;
	switch ( s.type )
	{
		case b2_circleShape:
			shape->circle = s.circle;
			break;
		case b2_capsuleShape:
			shape->capsule = s.capsule;
			break;
		case b2_polygonShape:
			shape->polygon = s.polygon;
			break;
		case b2_segmentShape:
			shape->segment = s.segment;
			break;
		default:
			break;
	}
	if ( bodyId != nullptr )
	{
		*bodyId = s.bodyId;
	}
	if ( friction != nullptr )
	{
		*friction = s.friction;
	}
	return 0;
}

int b2amd_Tape_Backward( b2amd_Tape* tape, const b2amd_BodyState* finalGrad, b2amd_BodyState* initialGrad, float* shapeFrictionGrad,
						 b2amd_AdjointStats* stats )
{
	TapeHost* t = reinterpret_cast<TapeHost*>( tape );
	if ( t->store.finished == false )
	{
		return -1;
	}
	int n = (int)t->store.worlds.size();
	std::vector<AdjointWorldArgs> args;
	std::vector<int> worldOf;
	size_t pointerCount = 0;
	for ( int i = 0; i < n; ++i )
	{
		const TapeStore::WorldTape& wt = t->store.worlds[i];
		if ( wt.steps.empty() || wt.final == nullptr || WorldOf( t->worldKeys[i] ) == nullptr )
		{
			continue;
		}
		pointerCount += wt.steps.size();
		worldOf.push_back( i );
	}
	int m = (int)worldOf.size();
	int bodyStride = std::max( t->bodyStride, 1 );
	int shapeStride = std::max( t->shapeStride, 1 );

	// scratch sizes
	size_t scratchFloats = 0;
	int ldsCap = 1;
	std::vector<size_t> scratchOff( m );
	std::vector<int> ncMax( m ), nbMax( m ), subMax( m );
	for ( int k = 0; k < m; ++k )
	{
		const TapeStore::WorldTape& wt = t->store.worlds[worldOf[k]];
		int nc = 1, nbm = std::max( 1, wt.finalBodyCount ), sm = 1;
		for ( size_t s = 0; s < wt.steps.size(); ++s )
		{
			nc = std::max( nc, wt.constraintCounts[s] );
			nbm = std::max( nbm, wt.bodyCounts[s] );
			sm = std::max( sm, wt.subStepCounts[s] );
		}
		ncMax[k] = nc;
		nbMax[k] = nbm;
		subMax[k] = sm;
		ldsCap = std::max( ldsCap, nbm + 1 );
		scratchOff[k] = scratchFloats;
		size_t floats = (size_t)sm * 3 * kOpFieldCount * nc + (size_t)sm * 3 * nbm + (size_t)kcFieldCount * nc + 6 * (size_t)nc +
						(size_t)kaFieldCount * nc + (size_t)kAccFieldCount * t->contactCaps[worldOf[k]];
		scratchFloats += ( floats + 63 ) & ~(size_t)63;
	}

	float* scratch = nullptr;
	char** devSteps = nullptr;
	b2amd_BodyState* devFinal = nullptr;
	b2amd_BodyState* devInitial = nullptr;
	float* devFriction = nullptr;
	int* devDiag = nullptr;
	AdjointWorldArgs* devArgs = nullptr;
	size_t gradCount = (size_t)n * bodyStride;
	HIP_CHECK( hipMalloc( &scratch, std::max<size_t>( scratchFloats, 1 ) * sizeof( float ) ) );
	HIP_CHECK( hipMalloc( &devSteps, std::max<size_t>( pointerCount, 1 ) * sizeof( char* ) ) );
	HIP_CHECK( hipMalloc( &devFinal, gradCount * sizeof( b2amd_BodyState ) ) );
	HIP_CHECK( hipMalloc( &devInitial, gradCount * sizeof( b2amd_BodyState ) ) );
	HIP_CHECK( hipMalloc( &devFriction, (size_t)n * shapeStride * sizeof( float ) ) );
	HIP_CHECK( hipMalloc( &devDiag, 2 * sizeof( int ) ) );
	HIP_CHECK( hipMalloc( &devArgs, std::max( m, 1 ) * sizeof( AdjointWorldArgs ) ) );

	std::vector<char*> stepPtrs;
	stepPtrs.reserve( pointerCount );
	for ( int k = 0; k < m; ++k )
	{
		int i = worldOf[k];
		const TapeStore::WorldTape& wt = t->store.worlds[i];
		AdjointWorldArgs a;
		a.steps = devSteps + stepPtrs.size();
		for ( char* r : wt.steps )
		{
			stepPtrs.push_back( r );
		}
		a.final = wt.final;
		a.stepCount = (int)wt.steps.size();
		a.nbMax = nbMax[k];
		a.ncMax = ncMax[k];
		a.contactCap = t->contactCaps[i];
		a.subStepMax = subMax[k];
		a.shapes = t->shapes[i];
		a.sims = t->sims[i];
		float* base = scratch + scratchOff[k];
		size_t nc = (size_t)ncMax[k];
		a.opTape = base;
		base += (size_t)subMax[k] * 3 * kOpFieldCount * nc + (size_t)subMax[k] * 3 * nbMax[k];
		a.cfields = base;
		base += (size_t)kcFieldCount * nc;
		a.cints = reinterpret_cast<int*>( base );
		base += 6 * nc;
		a.cadj = base;
		base += (size_t)kaFieldCount * nc;
		a.contactAcc = base;
		a.finalGrad = devFinal + (size_t)i * bodyStride;
		a.initialGrad = devInitial + (size_t)i * bodyStride;
		a.frictionGrad = devFriction + (size_t)i * shapeStride;
		a.diag = devDiag;
		args.push_back( a );
	}

	hipStream_t stream;
	HIP_CHECK( hipStreamCreate( &stream ) );
	hipEvent_t ev0, ev1;
	HIP_CHECK( hipEventCreate( &ev0 ) );
	HIP_CHECK( hipEventCreate( &ev1 ) );
	HIP_CHECK( hipMemcpyAsync( devSteps, stepPtrs.data(), stepPtrs.size() * sizeof( char* ), hipMemcpyHostToDevice, stream ) );
	HIP_CHECK( hipMemcpyAsync( devFinal, finalGrad, gradCount * sizeof( b2amd_BodyState ), hipMemcpyHostToDevice, stream ) );
	HIP_CHECK( hipMemsetAsync( devInitial, 0, gradCount * sizeof( b2amd_BodyState ), stream ) );
	HIP_CHECK( hipMemsetAsync( devFriction, 0, (size_t)n * shapeStride * sizeof( float ), stream ) );
	HIP_CHECK( hipMemsetAsync( devDiag, 0, 2 * sizeof( int ), stream ) );
	if ( m > 0 )
	{
		HIP_CHECK( hipMemcpyAsync( devArgs, args.data(), m * sizeof( AdjointWorldArgs ), hipMemcpyHostToDevice, stream ) );
	}

	size_t lds = (size_t)ldsCap * ( 18 * sizeof( float ) + sizeof( uint16_t ) );
	lds = ( lds + 15 ) & ~(size_t)15;
	if ( lds > 64 * 1024 )
	{
		HIP_CHECK( hipFuncSetAttribute( reinterpret_cast<const void*>( AdjointKernel ), hipFuncAttributeMaxDynamicSharedMemorySize,
										(int)lds ) );
	}
	HIP_CHECK( hipEventRecord( ev0, stream ) );
	if ( m > 0 )
	{
		AdjointKernel<<<m, kAdjointThreads, lds, stream>>>( devArgs );
		HIP_CHECK( hipGetLastError() );
	}
	HIP_CHECK( hipEventRecord( ev1, stream ) );

	int diag[2] = { 0, 0 };
	HIP_CHECK( hipMemcpyAsync( initialGrad, devInitial, gradCount * sizeof( b2amd_BodyState ), hipMemcpyDeviceToHost, stream ) );
	if ( shapeFrictionGrad != nullptr )
	{
		HIP_CHECK( hipMemcpyAsync( shapeFrictionGrad, devFriction, (size_t)n * shapeStride * sizeof( float ), hipMemcpyDeviceToHost,
								   stream ) );
	}
	HIP_CHECK( hipMemcpyAsync( diag, devDiag, sizeof( diag ), hipMemcpyDeviceToHost, stream ) );
	HIP_CHECK( hipStreamSynchronize( stream ) );
	float ms = 0.0f;
	HIP_CHECK( hipEventElapsedTime( &ms, ev0, ev1 ) );
	HIP_CHECK( hipEventDestroy( ev0 ) );
	HIP_CHECK( hipEventDestroy( ev1 ) );
	HIP_CHECK( hipStreamDestroy( stream ) );

	hipFree( scratch );
	hipFree( devSteps );
	hipFree( devFinal );
	hipFree( devInitial );
	hipFree( devFriction );
	hipFree( devDiag );
	hipFree( devArgs );

	if ( stats != nullptr )
	{
		stats->backwardMs = ms;
		stats->replayMismatches = diag[0];
		stats->unsupported = diag[1];
		stats->worlds = m;
		stats->steps = m > 0 ? (int)t->store.worlds[worldOf[0]].steps.size() : 0;
	}
	return 0;
}

} // extern "C"
