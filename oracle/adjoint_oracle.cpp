// CPU restatement of recorded Box2D steps for checking the GPU adjoint.
//
// TEST INFRASTRUCTURE ONLY. Loaded by tests/ to check b2amd_Tape_Backward; nothing in the product
// links or calls it.
//
// The replay re-runs the steps of one recorded world with the discrete structure of the record
// (contacts in solve order by graph color, recycled manifolds, warm-start point matches) and its
// own arithmetic, templated on the scalar type:
//   float   the replay must reproduce every recorded step bit for bit. This pins the restatement
//           (and the record) to the GPU forward, which is itself hash-identical to the reference.
//   double  a float64 model of the same steps, for central finite differences.
//   Rev     float64 reverse-mode automatic differentiation through a value tape: an adjoint
//           derived mechanically, independent of the hand-written GPU adjoint.
// Clamp and bias branches are taken by each scalar type on its own values; a decision signature
// (hash of every branch taken and every manifold feature id) tells whether two runs took the same
// branches, i.e. whether a finite difference stayed within one smooth piece.
//
// Reference code restated (operation order kept so float results match):
//   polygon manifold            src/manifold.c:511-1028, segment distance src/distance.c:33-106
//   manifold update, recycling  src/contact.c:519-698, src/physics_world.c:507-577
//   constraint preparation      src/contact_solver.c:1573-1809
//   warm start, solve, relax    src/contact_solver.c:1811-2116 (wide lanes: a >= b ? a : b max)
//   body integration, finalize  src/solver.c:66-162, 553-600
//   math helpers                include/box2d/math_functions.h

#include "box2d_amd.h"

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

namespace
{

// ---------------------------------------------------------------------------------------------
// Scalars

struct RevTape
{
	std::vector<int> a, b;
	std::vector<double> da, db;
	void Clear()
	{
		a.clear();
		b.clear();
		da.clear();
		db.clear();
	}
};

RevTape* g_tape = nullptr;
double* g_stepGrads = nullptr;

struct Rev
{
	double v;
	int i; // tape node, -1 for a constant
};

inline Rev Node( double v, int a, double da, int b, double db )
{
	if ( a < 0 && b < 0 )
	{
		return Rev{ v, -1 };
	}
	g_tape->a.push_back( a );
	g_tape->b.push_back( b );
	g_tape->da.push_back( da );
	g_tape->db.push_back( db );
	return Rev{ v, (int)g_tape->a.size() - 1 };
}

inline Rev operator+( Rev x, Rev y )
{
	return Node( x.v + y.v, x.i, 1.0, y.i, 1.0 );
}
inline Rev operator-( Rev x, Rev y )
{
	return Node( x.v - y.v, x.i, 1.0, y.i, -1.0 );
}
inline Rev operator*( Rev x, Rev y )
{
	return Node( x.v * y.v, x.i, y.v, y.i, x.v );
}
inline Rev operator/( Rev x, Rev y )
{
	double r = x.v / y.v;
	return Node( r, x.i, 1.0 / y.v, y.i, -r / y.v );
}
inline Rev operator-( Rev x )
{
	return Node( -x.v, x.i, -1.0, -1, 0.0 );
}
inline Rev Sqrt( Rev x )
{
	double r = std::sqrt( x.v );
	return Node( r, x.i, r > 0.0 ? 0.5 / r : 0.0, -1, 0.0 );
}
inline float Sqrt( float x )
{
	return sqrtf( x );
}
inline double Sqrt( double x )
{
	return std::sqrt( x );
}

inline double Val( float x )
{
	return x;
}
inline double Val( double x )
{
	return x;
}
inline double Val( Rev x )
{
	return x.v;
}

template <typename T>
inline T K( float x );
template <>
inline float K<float>( float x )
{
	return x;
}
template <>
inline double K<double>( float x )
{
	return (double)x;
}
template <>
inline Rev K<Rev>( float x )
{
	return Rev{ (double)x, -1 };
}

// Every branch the replay takes on a computed value goes through Dec. Recording keeps the branch
// pattern of a run (the float replay's is the recorded forward's); forcing makes a float64 or
// reverse-mode replay take that pattern, i.e. differentiate the recorded forward's smooth piece
// with the active set and clamp pattern held fixed. flips counts forced branches that the run's
// own values would have taken the other way.
struct Decisions
{
	std::vector<uint8_t> bits;
	size_t pos = 0;
	int mode = 0; // 0 off, 1 record, 2 force
	int flips = 0;
	bool overrun = false;
};

Decisions* g_dec = nullptr;

inline bool Dec( bool natural )
{
	if ( g_dec == nullptr || g_dec->mode == 0 )
	{
		return natural;
	}
	if ( g_dec->mode == 1 )
	{
		g_dec->bits.push_back( natural ? 1 : 0 );
		return natural;
	}
	if ( g_dec->pos >= g_dec->bits.size() )
	{
		g_dec->overrun = true;
		return natural;
	}
	bool b = g_dec->bits[g_dec->pos++] != 0;
	if ( b != natural )
	{
		g_dec->flips += 1;
	}
	return b;
}

// ---------------------------------------------------------------------------------------------
// Vectors and rotations (include/box2d/math_functions.h)

template <typename T>
struct V2
{
	T x, y;
};

template <typename T>
struct R2
{
	T c, s;
};

template <typename T>
inline T FromD( double x );
template <>
inline float FromD<float>( double x )
{
	return (float)x;
}
template <>
inline double FromD<double>( double x )
{
	return x;
}
template <>
inline Rev FromD<Rev>( double x )
{
	return Rev{ x, -1 };
}

template <typename T>
inline V2<T> Vk( b2Vec2 v )
{
	return V2<T>{ K<T>( v.x ), K<T>( v.y ) };
}

template <typename T>
inline V2<T> Add( V2<T> a, V2<T> b )
{
	return V2<T>{ a.x + b.x, a.y + b.y };
}
template <typename T>
inline V2<T> Sub( V2<T> a, V2<T> b )
{
	return V2<T>{ a.x - b.x, a.y - b.y };
}
template <typename T>
inline T Dot( V2<T> a, V2<T> b )
{
	return a.x * b.x + a.y * b.y;
}
template <typename T>
inline T Cross( V2<T> a, V2<T> b )
{
	return a.x * b.y - a.y * b.x;
}
template <typename T>
inline V2<T> MulAdd( V2<T> a, T s, V2<T> b )
{
	return V2<T>{ a.x + s * b.x, a.y + s * b.y };
}
template <typename T>
inline V2<T> Lerp( V2<T> a, V2<T> b, T t )
{
	T omt = K<T>( 1.0f ) - t;
	return V2<T>{ omt * a.x + t * b.x, omt * a.y + t * b.y };
}
template <typename T>
inline V2<T> Rotate( R2<T> q, V2<T> v )
{
	return V2<T>{ q.c * v.x - q.s * v.y, q.s * v.x + q.c * v.y };
}
template <typename T>
inline V2<T> InvRotate( R2<T> q, V2<T> v )
{
	return V2<T>{ q.c * v.x + q.s * v.y, -q.s * v.x + q.c * v.y };
}
template <typename T>
inline R2<T> MulRot( R2<T> q, R2<T> r )
{
	R2<T> qr;
	qr.s = q.s * r.c + q.c * r.s;
	qr.c = q.c * r.c - q.s * r.s;
	return qr;
}
template <typename T>
inline R2<T> InvMulRot( R2<T> a, R2<T> b )
{
	R2<T> r;
	r.s = a.c * b.s - a.s * b.c;
	r.c = a.c * b.c + a.s * b.s;
	return r;
}
template <typename T>
inline R2<T> Invert( R2<T> a )
{
	return R2<T>{ a.c, -a.s };
}
template <typename T>
inline R2<T> NormalizeRot( R2<T> q )
{
	T mag = Sqrt( q.s * q.s + q.c * q.c );
	T invMag = Dec( Val( mag ) > 0.0 ) ? K<T>( 1.0f ) / mag : K<T>( 0.0f );
	return R2<T>{ q.c * invMag, q.s * invMag };
}
template <typename T>
inline R2<T> IntegrateRotation( R2<T> q1, T deltaAngle )
{
	R2<T> q2 = { q1.c - deltaAngle * q1.s, q1.s + deltaAngle * q1.c };
	T mag = Sqrt( q2.s * q2.s + q2.c * q2.c );
	T invMag = Dec( Val( mag ) > 0.0 ) ? K<T>( 1.0f ) / mag : K<T>( 0.0f );
	return R2<T>{ q2.c * invMag, q2.s * invMag };
}
template <typename T>
inline T Clamp01( T a )
{
	if ( Dec( Val( a ) < 0.0 ) )
	{
		return K<T>( 0.0f );
	}
	return Dec( Val( a ) > 1.0 ) ? K<T>( 1.0f ) : a;
}

// Branch and feature record of a run (FNV-1a over 32-bit words).
struct Signature
{
	uint64_t h = 1469598103934665603ull;
	void Add( uint32_t w )
	{
		h ^= w;
		h *= 1099511628211ull;
	}
};

// ---------------------------------------------------------------------------------------------
// Polygon manifold (src/manifold.c:511-1028)

constexpr float kLinearSlop = 0.005f;
constexpr float kSpeculative = 4.0f * kLinearSlop;

template <typename T>
struct Poly
{
	V2<T> v[8];
	V2<T> n[8];
	int count;
	float radius;
};

template <typename T>
struct LocalPoint
{
	V2<T> point;
	T separation;
	uint16_t id;
};

template <typename T>
struct Local
{
	V2<T> normal;
	LocalPoint<T> p[2];
	int count;
};

inline uint16_t MakeId( int a, int b )
{
	return (uint16_t)( ( (uint8_t)a << 8 ) | (uint8_t)b );
}

template <typename T>
Local<T> EmptyLocal()
{
	Local<T> m;
	m.normal = Vk<T>( { 0.0f, 0.0f } );
	for ( int j = 0; j < 2; ++j )
	{
		m.p[j].point = Vk<T>( { 0.0f, 0.0f } );
		m.p[j].separation = K<T>( 0.0f );
		m.p[j].id = 0;
	}
	m.count = 0;
	return m;
}

template <typename T>
double MaxSeparation( int* edge, const Poly<T>& p1, const Poly<T>& p2 )
{
	int best = 0;
	double maxSep = -FLT_MAX;
	for ( int i = 0; i < p1.count; ++i )
	{
		double si = FLT_MAX;
		for ( int j = 0; j < p2.count; ++j )
		{
			double sij = Val( Dot( p1.n[i], Sub( p2.v[j], p1.v[i] ) ) );
			if ( Dec( sij < si ) )
			{
				si = sij;
			}
		}
		if ( Dec( si > maxSep ) )
		{
			maxSep = si;
			best = i;
		}
	}
	*edge = best;
	return maxSep;
}

template <typename T>
Local<T> Clip( const Poly<T>& polyA, const Poly<T>& polyB, int edgeA, int edgeB, bool flip )
{
	Local<T> m = EmptyLocal<T>();
	const Poly<T>* p1 = flip ? &polyB : &polyA;
	const Poly<T>* p2 = flip ? &polyA : &polyB;
	int i11 = flip ? edgeB : edgeA;
	int i12 = i11 + 1 < p1->count ? i11 + 1 : 0;
	int i21 = flip ? edgeA : edgeB;
	int i22 = i21 + 1 < p2->count ? i21 + 1 : 0;

	V2<T> normal = p1->n[i11];
	V2<T> v11 = p1->v[i11], v12 = p1->v[i12], v21 = p2->v[i21], v22 = p2->v[i22];
	V2<T> tangent = { -K<T>( 1.0f ) * normal.y, K<T>( 1.0f ) * normal.x };

	T lower1 = K<T>( 0.0f );
	T upper1 = Dot( Sub( v12, v11 ), tangent );
	T upper2 = Dot( Sub( v21, v11 ), tangent );
	T lower2 = Dot( Sub( v22, v11 ), tangent );
	if ( Dec( Val( upper2 ) < Val( lower1 ) || Val( upper1 ) < Val( lower2 ) ) )
	{
		return m;
	}

	V2<T> vLower = v22;
	if ( Dec( Val( lower2 ) < Val( lower1 ) && Val( upper2 - lower2 ) > FLT_EPSILON ) )
	{
		vLower = Lerp( v22, v21, ( lower1 - lower2 ) / ( upper2 - lower2 ) );
	}
	V2<T> vUpper = v21;
	if ( Dec( Val( upper2 ) > Val( upper1 ) && Val( upper2 - lower2 ) > FLT_EPSILON ) )
	{
		vUpper = Lerp( v22, v21, ( upper1 - lower2 ) / ( upper2 - lower2 ) );
	}

	T sepLower = Dot( Sub( vLower, v11 ), normal );
	T sepUpper = Dot( Sub( vUpper, v11 ), normal );
	float r1 = p1->radius, r2 = p2->radius;
	vLower = MulAdd( vLower, K<T>( 0.5f ) * ( K<T>( r1 - r2 ) - sepLower ), normal );
	vUpper = MulAdd( vUpper, K<T>( 0.5f ) * ( K<T>( r1 - r2 ) - sepUpper ), normal );
	float radius = r1 + r2;

	if ( flip == false )
	{
		m.normal = normal;
		m.p[0] = { vLower, sepLower - K<T>( radius ), MakeId( i11, i22 ) };
		m.p[1] = { vUpper, sepUpper - K<T>( radius ), MakeId( i12, i21 ) };
	}
	else
	{
		m.normal = V2<T>{ -normal.x, -normal.y };
		m.p[0] = { vUpper, sepUpper - K<T>( radius ), MakeId( i21, i12 ) };
		m.p[1] = { vLower, sepLower - K<T>( radius ), MakeId( i22, i11 ) };
	}
	m.count = 2;
	return m;
}

template <typename T>
struct SegDist
{
	V2<T> c1, c2;
	T f1, f2, dd;
};

template <typename T>
SegDist<T> SegmentDistance( V2<T> p1, V2<T> q1, V2<T> p2, V2<T> q2 )
{
	SegDist<T> r;
	r.f1 = K<T>( 0.0f );
	r.f2 = K<T>( 0.0f );
	V2<T> d1 = Sub( q1, p1 ), d2 = Sub( q2, p2 ), rr = Sub( p1, p2 );
	T dd1 = Dot( d1, d1 ), dd2 = Dot( d2, d2 ), rd1 = Dot( rr, d1 ), rd2 = Dot( rr, d2 );
	const double epsSqr = (double)( FLT_EPSILON * FLT_EPSILON );
	if ( Dec( Val( dd1 ) < epsSqr || Val( dd2 ) < epsSqr ) )
	{
		if ( Dec( Val( dd1 ) >= epsSqr ) )
		{
			r.f1 = Clamp01( -rd1 / dd1 );
		}
		else if ( Dec( Val( dd2 ) >= epsSqr ) )
		{
			r.f2 = Clamp01( rd2 / dd2 );
		}
	}
	else
	{
		T d12 = Dot( d1, d2 );
		T den = dd1 * dd2 - d12 * d12;
		T f1 = K<T>( 0.0f );
		if ( Dec( Val( den ) != 0.0 ) )
		{
			f1 = Clamp01( ( d12 * rd2 - rd1 * dd2 ) / den );
		}
		T f2 = ( d12 * f1 + rd2 ) / dd2;
		if ( Dec( Val( f2 ) < 0.0 ) )
		{
			f2 = K<T>( 0.0f );
			f1 = Clamp01( -rd1 / dd1 );
		}
		else if ( Dec( Val( f2 ) > 1.0 ) )
		{
			f2 = K<T>( 1.0f );
			f1 = Clamp01( ( d12 - rd1 ) / dd1 );
		}
		r.f1 = f1;
		r.f2 = f2;
	}
	r.c1 = MulAdd( p1, r.f1, d1 );
	r.c2 = MulAdd( p2, r.f2, d2 );
	V2<T> c = Sub( r.c2, r.c1 );
	r.dd = c.x * c.x + c.y * c.y;
	return r;
}

// Manifold of polygon B in polygon A's frame, xf the pose of B relative to A.
template <typename T>
Local<T> CollidePolygons( const b2Polygon& A, const b2Polygon& B, V2<T> xp, R2<T> xq )
{
	b2Vec2 origin = A.vertices[0];
	V2<T> sp = { xp.x - K<T>( origin.x ), xp.y - K<T>( origin.y ) };
	Poly<T> pa, pb;
	pa.count = A.count;
	pa.radius = A.radius;
	pa.v[0] = Vk<T>( { 0.0f, 0.0f } );
	pa.n[0] = Vk<T>( A.normals[0] );
	for ( int i = 1; i < A.count; ++i )
	{
		pa.v[i] = Vk<T>( { A.vertices[i].x - origin.x, A.vertices[i].y - origin.y } );
		pa.n[i] = Vk<T>( A.normals[i] );
	}
	pb.count = B.count;
	pb.radius = B.radius;
	for ( int i = 0; i < B.count; ++i )
	{
		V2<T> v = Vk<T>( B.vertices[i] );
		pb.v[i] = V2<T>{ ( xq.c * v.x - xq.s * v.y ) + sp.x, ( xq.s * v.x + xq.c * v.y ) + sp.y };
		pb.n[i] = Rotate( xq, Vk<T>( B.normals[i] ) );
	}

	int edgeA = 0, edgeB = 0;
	double sepA = MaxSeparation( &edgeA, pa, pb );
	double sepB = MaxSeparation( &edgeB, pb, pa );
	float radius = pa.radius + pb.radius;
	if ( Dec( sepA > (double)( kSpeculative + radius ) || sepB > (double)( kSpeculative + radius ) ) )
	{
		return EmptyLocal<T>();
	}

	bool flip = !Dec( sepA >= sepB );
	if ( flip == false )
	{
		edgeB = 0;
		double minDot = FLT_MAX;
		for ( int i = 0; i < pb.count; ++i )
		{
			double d = Val( Dot( pa.n[edgeA], pb.n[i] ) );
			if ( Dec( d < minDot ) )
			{
				minDot = d;
				edgeB = i;
			}
		}
	}
	else
	{
		edgeA = 0;
		double minDot = FLT_MAX;
		for ( int i = 0; i < pa.count; ++i )
		{
			double d = Val( Dot( pb.n[edgeB], pa.n[i] ) );
			if ( Dec( d < minDot ) )
			{
				minDot = d;
				edgeA = i;
			}
		}
	}

	Local<T> m = EmptyLocal<T>();
	if ( Dec( sepA > (double)( 0.1f * kLinearSlop ) || sepB > (double)( 0.1f * kLinearSlop ) ) )
	{
		int i11 = edgeA, i12 = edgeA + 1 < pa.count ? edgeA + 1 : 0;
		int i21 = edgeB, i22 = edgeB + 1 < pb.count ? edgeB + 1 : 0;
		V2<T> v11 = pa.v[i11], v12 = pa.v[i12], v21 = pb.v[i21], v22 = pb.v[i22];
		SegDist<T> r = SegmentDistance( v11, v12, v21, v22 );
		T distance = Sqrt( r.dd );
		T separation = distance - K<T>( radius );
		if ( Dec( Val( distance - K<T>( radius ) ) > (double)kSpeculative ) )
		{
			return m;
		}
		m = Clip( pa, pb, edgeA, edgeB, flip );
		double minSep = FLT_MAX;
		for ( int i = 0; i < m.count; ++i )
		{
			minSep = Dec( Val( m.p[i].separation ) < minSep ) ? Val( m.p[i].separation ) : minSep;
		}
		if ( Dec( Val( separation + K<T>( 0.1f * kLinearSlop ) ) < minSep ) )
		{
			double f1 = Val( r.f1 ), f2 = Val( r.f2 );
			int iA = -1, iB = -1;
			V2<T> vA = v11, vB = v21;
			if ( Dec( f1 == 0.0 && f2 == 0.0 ) )
			{
				iA = i11, iB = i21, vA = v11, vB = v21;
			}
			else if ( Dec( f1 == 0.0 && f2 == 1.0 ) )
			{
				iA = i11, iB = i22, vA = v11, vB = v22;
			}
			else if ( Dec( f1 == 1.0 && f2 == 0.0 ) )
			{
				iA = i12, iB = i21, vA = v12, vB = v21;
			}
			else if ( Dec( f1 == 1.0 && f2 == 1.0 ) )
			{
				iA = i12, iB = i22, vA = v12, vB = v22;
			}
			if ( iA >= 0 )
			{
				V2<T> normal = Sub( vB, vA );
				T inv = K<T>( 1.0f ) / distance;
				normal.x = normal.x * inv;
				normal.y = normal.y * inv;
				V2<T> c1 = MulAdd( vA, K<T>( pa.radius ), normal );
				V2<T> c2 = MulAdd( vB, K<T>( -pb.radius ), normal );
				m.normal = normal;
				m.p[0] = { Lerp( c1, c2, K<T>( 0.5f ) ), distance - K<T>( radius ), MakeId( iA, iB ) };
				m.count = 1;
			}
		}
	}
	else
	{
		m = Clip( pa, pb, edgeA, edgeB, flip );
	}
	for ( int i = 0; i < m.count; ++i )
	{
		m.p[i].point = Add( m.p[i].point, Vk<T>( origin ) );
	}
	return m;
}

// ---------------------------------------------------------------------------------------------
// Replay state

template <typename T>
struct Body
{
	V2<T> c;
	R2<T> q;
	V2<T> v;
	T w;
	V2<T> dp;
	R2<T> dq;
	float invMass, invI, gravityScale, linearDamping, angularDamping, torque;
	b2Vec2 localCenter, force;
	bool dynamic;
};

template <typename T>
struct ContactState
{
	V2<T> normal;
	V2<T> anchorA[2], anchorB[2];
	T base[2];
	R2<T> cqA, cqB;
	T normalImpulse[2], tangentImpulse[2];
	int count = 0;
	uint16_t id[2];
};

template <typename T>
struct Constraint
{
	int a, b; // slots, 0 static
	float mA, mB, iA, iB;
	V2<T> n;
	V2<T> rA[2], rB[2];
	T base[2], nm[2], tm[2];
	T ln[2], lt[2];
	T friction;
	float biasRate, massScale, impulseScale, tangentSpeed;
	int count;
};

struct Input
{
	int stepCount;
	const b2amd_TapeStep* headers; // stepCount + 1
	const b2amd_TapeBody* bodies;
	const int* bodyStart; // stepCount + 1
	const b2amd_TapeConstraint* cons;
	const int* consStart; // stepCount
	const b2amd_Shape* shapes;
	const float* shapeFriction;
	int shapeCount;
	const b2amd_TapeBody* statics; // by body id (only static bodies used)
	int staticCount;
};

inline bool SameBits( float a, float b )
{
	uint32_t x, y;
	memcpy( &x, &a, 4 );
	memcpy( &y, &b, 4 );
	return x == y;
}

template <typename T>
struct Replay
{
	const Input& in;
	std::vector<Body<T>> bodies; // slot 0 static identity
	std::map<int, ContactState<T>> contacts;
	std::vector<T> friction; // per shape
	Signature sig;
	int floatMismatch = 0;
	int structureMismatch = 0;

	explicit Replay( const Input& input )
		: in( input )
	{
	}

	float Fm( T x )
	{
		return (float)Val( x );
	}

	void CheckFloat( bool same )
	{
		if ( same == false )
		{
			floatMismatch += 1;
		}
	}

	// pose of a constraint side: center, origin, rotation, local center
	void PoseOf( int slot, int bodyId, V2<T>* c, V2<T>* p, R2<T>* q, b2Vec2* lc )
	{
		if ( slot != 0 )
		{
			const Body<T>& b = bodies[slot];
			*c = b.c;
			*q = b.q;
			*lc = b.localCenter;
			// origin = center - R(q) localCenter (finalize's b2OffsetPos( center, -R(q) lc ))
			V2<T> off = Rotate( b.q, Vk<T>( b.localCenter ) );
			*p = V2<T>{ b.c.x + ( -off.x ), b.c.y + ( -off.y ) };
		}
		else
		{
			const b2amd_TapeBody& s = in.statics[bodyId];
			*c = Vk<T>( s.center );
			*p = Vk<T>( s.origin );
			*q = R2<T>{ K<T>( s.rotation.c ), K<T>( s.rotation.s ) };
			*lc = s.localCenter;
		}
	}

	void Step( int k )
	{
		const b2amd_TapeStep& h = in.headers[k];
		const b2amd_TapeConstraint* rc = in.cons + in.consStart[k];
		const b2amd_TapeBody* rb = in.bodies + in.bodyStart[k];
		int nb = h.bodyCount, nc = h.constraintCount;
		const bool floatMode = sizeof( T ) == sizeof( float );

		// constraints from the manifolds
		std::vector<Constraint<T>> cons( nc );
		for ( int i = 0; i < nc; ++i )
		{
			const b2amd_TapeConstraint& r = rc[i];
			Constraint<T>& c = cons[i];
			c.a = r.slotA;
			c.b = r.slotB;
			c.count = r.pointCount;
			ContactState<T>& cs = contacts[r.contactId];
			V2<T> cA, pA, cB, pB;
			R2<T> qA, qB;
			b2Vec2 lcA, lcB;
			PoseOf( r.slotA, r.bodyIdA, &cA, &pA, &qA, &lcA );
			PoseOf( r.slotB, r.bodyIdB, &cB, &pB, &qB, &lcB );

			T sep[2] = { K<T>( 0.0f ), K<T>( 0.0f ) };
			if ( r.info & 1u )
			{
				// recycled: keep anchors and normal, update separations
				R2<T> dqA = MulRot( qA, Invert( cs.cqA ) );
				R2<T> dqB = MulRot( qB, Invert( cs.cqB ) );
				V2<T> dc = Sub( cB, cA );
				for ( int j = 0; j < cs.count; ++j )
				{
					V2<T> ra = Rotate( dqA, cs.anchorA[j] );
					V2<T> rb2 = Rotate( dqB, cs.anchorB[j] );
					V2<T> d = Add( dc, Sub( rb2, ra ) );
					sep[j] = cs.base[j] + Dot( d, cs.normal );
				}
				if ( cs.count != r.pointCount )
				{
					structureMismatch += 1;
				}
			}
			else
			{
				// full update: relative transform, polygon manifold, world anchors
				ContactState<T> old = cs;
				R2<T> rq = InvMulRot( qA, qB );
				V2<T> d = { pB.x - pA.x, pB.y - pA.y };
				V2<T> rp = InvRotate( qA, d );
				const b2amd_Shape& sA = in.shapes[r.shapeIdA];
				const b2amd_Shape& sB = in.shapes[r.shapeIdB];
				Local<T> m = CollidePolygons<T>( sA.polygon, sB.polygon, rp, rq );
				if ( m.count != r.pointCount )
				{
					structureMismatch += 1;
				}
				cs.count = m.count < r.pointCount ? m.count : r.pointCount;
				cs.normal = Rotate( qA, m.normal );
				V2<T> originDelta = { pA.x - pB.x, pA.y - pB.y };
				V2<T> coA = Rotate( qA, Vk<T>( lcA ) );
				V2<T> coB = Rotate( qB, Vk<T>( lcB ) );
				for ( int j = 0; j < cs.count; ++j )
				{
					if ( m.p[j].id != r.id[j] )
					{
						structureMismatch += 1;
					}
					V2<T> aw = Rotate( qA, m.p[j].point );
					V2<T> bw = Add( aw, originDelta );
					cs.anchorA[j] = Sub( aw, coA );
					cs.anchorB[j] = Sub( bw, coB );
					cs.base[j] = m.p[j].separation;
					sep[j] = m.p[j].separation;
					cs.id[j] = m.p[j].id;
					sig.Add( m.p[j].id );
					int src = ( r.info >> ( 2 + 2 * j ) ) & 3;
					cs.normalImpulse[j] = src > 0 ? old.normalImpulse[src - 1] : K<T>( 0.0f );
					cs.tangentImpulse[j] = src > 0 ? old.tangentImpulse[src - 1] : K<T>( 0.0f );
				}
				cs.cqA = qA;
				cs.cqB = qB;
			}

			// friction sqrt(fA fB), constant over the run
			T fA = friction[r.shapeIdA], fB = friction[r.shapeIdB];
			c.friction = Sqrt( fA * fB );

			if ( floatMode )
			{
				bool same = SameBits( Fm( cs.normal.x ), r.normal.x ) && SameBits( Fm( cs.normal.y ), r.normal.y ) &&
							SameBits( Fm( c.friction ), r.friction );
				for ( int j = 0; j < r.pointCount && j < cs.count; ++j )
				{
					same = same && SameBits( Fm( cs.anchorA[j].x ), r.anchorA[j].x ) && SameBits( Fm( cs.anchorA[j].y ), r.anchorA[j].y ) &&
						   SameBits( Fm( cs.anchorB[j].x ), r.anchorB[j].x ) && SameBits( Fm( cs.anchorB[j].y ), r.anchorB[j].y ) &&
						   SameBits( Fm( sep[j] ), r.separation[j] ) && SameBits( Fm( cs.normalImpulse[j] ), r.normalImpulse[j] ) &&
						   SameBits( Fm( cs.tangentImpulse[j] ), r.tangentImpulse[j] );
				}
				CheckFloat( same );
			}

			// preparation (src/contact_solver.c:1573-1809)
			c.mA = r.slotA != 0 ? bodies[r.slotA].invMass : 0.0f;
			c.iA = r.slotA != 0 ? bodies[r.slotA].invI : 0.0f;
			c.mB = r.slotB != 0 ? bodies[r.slotB].invMass : 0.0f;
			c.iB = r.slotB != 0 ? bodies[r.slotB].invI : 0.0f;
			c.biasRate = r.biasRate;
			c.massScale = r.massScale;
			c.impulseScale = r.impulseScale;
			c.tangentSpeed = r.tangentSpeed;
			c.n = cs.normal;
			V2<T> t = { c.n.y, -c.n.x };
			T ws = K<T>( h.enableWarmStarting ? 1.0f : 0.0f );
			for ( int j = 0; j < 2; ++j )
			{
				if ( j < c.count )
				{
					c.rA[j] = cs.anchorA[j];
					c.rB[j] = cs.anchorB[j];
					c.base[j] = sep[j] - Dot( Sub( c.rB[j], c.rA[j] ), c.n );
					c.ln[j] = ws * cs.normalImpulse[j];
					c.lt[j] = ws * cs.tangentImpulse[j];
					T rnA = Cross( c.rA[j], c.n ), rnB = Cross( c.rB[j], c.n );
					T kN = K<T>( c.mA ) + K<T>( c.mB ) + K<T>( c.iA ) * rnA * rnA + K<T>( c.iB ) * rnB * rnB;
					c.nm[j] = Dec( Val( kN ) > 0.0 ) ? K<T>( 1.0f ) / kN : K<T>( 0.0f );
					T rtA = Cross( c.rA[j], t ), rtB = Cross( c.rB[j], t );
					T kT = K<T>( c.mA ) + K<T>( c.mB ) + K<T>( c.iA ) * rtA * rtA + K<T>( c.iB ) * rtB * rtB;
					c.tm[j] = Dec( Val( kT ) > 0.0 ) ? K<T>( 1.0f ) / kT : K<T>( 0.0f );
				}
				else
				{
					c.rA[j] = Vk<T>( { 0.0f, 0.0f } );
					c.rB[j] = Vk<T>( { 0.0f, 0.0f } );
					c.base[j] = K<T>( 0.0f );
					c.ln[j] = K<T>( 0.0f );
					c.lt[j] = K<T>( 0.0f );
					c.nm[j] = K<T>( 0.0f );
					c.tm[j] = K<T>( 0.0f );
				}
			}
		}

		// substeps (src/solver.c:1090-1170 stage order)
		T hh = K<T>( h.h );
		for ( int sub = 0; sub < h.subStepCount; ++sub )
		{
			for ( int s = 1; s <= nb; ++s )
			{
				Body<T>& b = bodies[s];
				// src/solver.c:66-112
				T linearDamping = K<T>( 1.0f ) / ( K<T>( 1.0f ) + hh * K<T>( b.linearDamping ) );
				T angularDamping = K<T>( 1.0f ) / ( K<T>( 1.0f ) + hh * K<T>( b.angularDamping ) );
				T gravityScale = K<T>( b.invMass > 0.0f ? b.gravityScale : 0.0f );
				T hm = hh * K<T>( b.invMass );
				T hg = hh * gravityScale;
				V2<T> lvd = { hm * K<T>( b.force.x ) + hg * K<T>( h.gravity.x ), hm * K<T>( b.force.y ) + hg * K<T>( h.gravity.y ) };
				T avd = hh * K<T>( b.invI ) * K<T>( b.torque );
				b.v = V2<T>{ lvd.x + linearDamping * b.v.x, lvd.y + linearDamping * b.v.y };
				b.w = avd + angularDamping * b.w;
			}
			ForColors( h, cons, 0 );
			ForColors( h, cons, 1 );
			for ( int s = 1; s <= nb; ++s )
			{
				Body<T>& b = bodies[s];
				// src/solver.c:114-162 (no locks; speed caps are not differentiable and not expected)
				b.dp = MulAdd( b.dp, hh, b.v );
				b.dq = IntegrateRotation( b.dq, hh * b.w );
			}
			ForColors( h, cons, 2 );
		}

		// store impulses
		for ( int i = 0; i < nc; ++i )
		{
			ContactState<T>& cs = contacts[rc[i].contactId];
			for ( int j = 0; j < cons[i].count; ++j )
			{
				cs.normalImpulse[j] = cons[i].ln[j];
				cs.tangentImpulse[j] = cons[i].lt[j];
			}
		}

		// finalize (src/solver.c:553-600)
		for ( int s = 1; s <= nb; ++s )
		{
			Body<T>& b = bodies[s];
			b.c = Add( b.c, b.dp );
			b.q = NormalizeRot( MulRot( b.dq, b.q ) );
			b.dp = Vk<T>( { 0.0f, 0.0f } );
			b.dq = R2<T>{ K<T>( 1.0f ), K<T>( 0.0f ) };
		}

		if ( floatMode )
		{
			const b2amd_TapeBody* nbd = in.bodies + in.bodyStart[k + 1];
			for ( int s = 1; s <= nb; ++s )
			{
				const Body<T>& b = bodies[s];
				const b2amd_TapeBody& r = nbd[s - 1];
				CheckFloat( SameBits( Fm( b.c.x ), r.center.x ) && SameBits( Fm( b.c.y ), r.center.y ) &&
							SameBits( Fm( b.q.c ), r.rotation.c ) && SameBits( Fm( b.q.s ), r.rotation.s ) &&
							SameBits( Fm( b.v.x ), r.linearVelocity.x ) && SameBits( Fm( b.v.y ), r.linearVelocity.y ) &&
							SameBits( Fm( b.w ), r.angularVelocity ) );
			}
		}
		(void)rb;
	}

	void ForColors( const b2amd_TapeStep& h, std::vector<Constraint<T>>& cons, int stage )
	{
		T inv_h = K<T>( h.inv_h );
		T contactSpeed = K<T>( -h.contactSpeed );
		for ( int color = 0; color < h.activeColorCount; ++color )
		{
			for ( int i = h.colorStart[color]; i < h.colorStart[color + 1]; ++i )
			{
				if ( stage == 0 )
				{
					WarmStart( cons[i] );
				}
				else
				{
					Solve( cons[i], stage == 1, inv_h, contactSpeed );
				}
			}
		}
	}

	// src/contact_solver.c:1811-1871 (wide lanes)
	void WarmStart( Constraint<T>& c )
	{
		Body<T> staticBody = Static();
		Body<T>& A = c.a != 0 ? bodies[c.a] : staticBody;
		Body<T>& B = c.b != 0 ? bodies[c.b] : staticBody;
		V2<T> t = { c.n.y, K<T>( 0.0f ) - c.n.x };
		// lane copies: both points update them, only dynamic bodies are written back
		Body<T> a = A, b = B;
		for ( int j = 0; j < 2; ++j )
		{
			T Px = c.ln[j] * c.n.x + c.lt[j] * t.x;
			T Py = c.ln[j] * c.n.y + c.lt[j] * t.y;
			a.w = a.w - K<T>( c.iA ) * ( c.rA[j].x * Py - c.rA[j].y * Px );
			a.v.x = a.v.x - K<T>( c.mA ) * Px;
			a.v.y = a.v.y - K<T>( c.mA ) * Py;
			b.w = b.w + K<T>( c.iB ) * ( c.rB[j].x * Py - c.rB[j].y * Px );
			b.v.x = b.v.x + K<T>( c.mB ) * Px;
			b.v.y = b.v.y + K<T>( c.mB ) * Py;
		}
		Store( A, a );
		Store( B, b );
	}

	Body<T> Static()
	{
		Body<T> s;
		s.c = Vk<T>( { 0.0f, 0.0f } );
		s.q = R2<T>{ K<T>( 1.0f ), K<T>( 0.0f ) };
		s.v = Vk<T>( { 0.0f, 0.0f } );
		s.w = K<T>( 0.0f );
		s.dp = Vk<T>( { 0.0f, 0.0f } );
		s.dq = R2<T>{ K<T>( 1.0f ), K<T>( 0.0f ) };
		s.dynamic = false;
		return s;
	}

	void Store( Body<T>& dst, const Body<T>& lane )
	{
		if ( dst.dynamic )
		{
			dst.v = lane.v;
			dst.w = lane.w;
		}
	}

	// src/contact_solver.c:1873-2116 (wide lanes), rolling resistance and restitution absent
	void Solve( Constraint<T>& c, bool useBias, T inv_h, T contactSpeed )
	{
		Body<T> staticBody = Static();
		Body<T>& A = c.a != 0 ? bodies[c.a] : staticBody;
		Body<T>& B = c.b != 0 ? bodies[c.b] : staticBody;
		Body<T> a = A, b = B;

		T biasRate = useBias ? K<T>( c.massScale ) * K<T>( c.biasRate ) : K<T>( 0.0f );
		T massScale = useBias ? K<T>( c.massScale ) : K<T>( 1.0f );
		T impulseScale = useBias ? K<T>( c.impulseScale ) : K<T>( 0.0f );
		T dpX = b.dp.x - a.dp.x, dpY = b.dp.y - a.dp.y;

		for ( int j = 0; j < 2; ++j )
		{
			V2<T> rA = c.rA[j], rB = c.rB[j];
			T rsAx = a.dq.c * rA.x - a.dq.s * rA.y;
			T rsAy = a.dq.s * rA.x + a.dq.c * rA.y;
			T rsBx = b.dq.c * rB.x - b.dq.s * rB.y;
			T rsBy = b.dq.s * rB.x + b.dq.c * rB.y;
			T dsx = dpX + ( rsBx - rsAx );
			T dsy = dpY + ( rsBy - rsAy );
			T s = ( c.n.x * dsx + c.n.y * dsy ) + c.base[j];
			bool mask = Dec( Val( s ) > 0.0 );
			T soft = biasRate * s;
			bool softActive = Dec( Val( soft ) >= Val( contactSpeed ) );
			T bias = mask ? s * inv_h : ( softActive ? soft : contactSpeed );
			T pms = mask ? K<T>( 1.0f ) : massScale;
			T pis = mask ? K<T>( 0.0f ) : impulseScale;
			T dvx = ( b.v.x - b.w * rB.y ) - ( a.v.x - a.w * rA.y );
			T dvy = ( b.v.y + b.w * rB.x ) - ( a.v.y + a.w * rA.x );
			T vn = dvx * c.n.x + dvy * c.n.y;
			T neg = c.nm[j] * ( pms * vn + bias ) + pis * c.ln[j];
			T cand = c.ln[j] - neg;
			bool active = Dec( Val( cand ) >= 0.0 );
			T newImpulse = active ? cand : K<T>( 0.0f );
			T impulse = newImpulse - c.ln[j];
			c.ln[j] = newImpulse;
			T Px = impulse * c.n.x, Py = impulse * c.n.y;
			a.v.x = a.v.x - K<T>( c.mA ) * Px;
			a.v.y = a.v.y - K<T>( c.mA ) * Py;
			a.w = a.w - K<T>( c.iA ) * ( rA.x * Py - rA.y * Px );
			b.v.x = b.v.x + K<T>( c.mB ) * Px;
			b.v.y = b.v.y + K<T>( c.mB ) * Py;
			b.w = b.w + K<T>( c.iB ) * ( rB.x * Py - rB.y * Px );
			if ( j < c.count )
			{
				sig.Add( ( mask ? 1u : 0u ) | ( softActive ? 2u : 0u ) | ( active ? 4u : 0u ) );
			}
		}

		if ( useBias == false )
		{
			V2<T> t = { c.n.y, K<T>( 0.0f ) - c.n.x };
			for ( int j = 0; j < 2; ++j )
			{
				V2<T> rA = c.rA[j], rB = c.rB[j];
				T dvx = ( b.v.x - b.w * rB.y ) - ( a.v.x - a.w * rA.y );
				T dvy = ( b.v.y + b.w * rB.x ) - ( a.v.y + a.w * rA.x );
				T vt = dvx * t.x + dvy * t.y;
				vt = vt - K<T>( c.tangentSpeed );
				T neg = c.tm[j] * vt;
				T maxF = c.friction * c.ln[j];
				T cand = c.lt[j] - neg;
				bool innerIsCand = Dec( Val( cand ) <= Val( maxF ) );
				T inner = innerIsCand ? cand : maxF;
				T lower = K<T>( 0.0f ) - maxF;
				bool resIsLower = Dec( Val( lower ) >= Val( inner ) );
				T newImpulse = resIsLower ? lower : inner;
				T impulse = newImpulse - c.lt[j];
				c.lt[j] = newImpulse;
				T Px = impulse * t.x, Py = impulse * t.y;
				a.v.x = a.v.x - K<T>( c.mA ) * Px;
				a.v.y = a.v.y - K<T>( c.mA ) * Py;
				a.w = a.w - K<T>( c.iA ) * ( rA.x * Py - rA.y * Px );
				b.v.x = b.v.x + K<T>( c.mB ) * Px;
				b.v.y = b.v.y + K<T>( c.mB ) * Py;
				b.w = b.w + K<T>( c.iB ) * ( rB.x * Py - rB.y * Px );
				if ( j < c.count )
				{
					sig.Add( 8u | ( innerIsCand ? 1u : 0u ) | ( resIsLower ? 2u : 0u ) );
				}
			}
		}
		Store( A, a );
		Store( B, b );
	}
};

template <typename T>
void InitBodies( Replay<T>& r, const double* initState )
{
	const b2amd_TapeStep& h = r.in.headers[0];
	const b2amd_TapeBody* rb = r.in.bodies + r.in.bodyStart[0];
	r.bodies.resize( h.bodyCount + 1 );
	r.bodies[0] = r.Static();
	for ( int s = 1; s <= h.bodyCount; ++s )
	{
		const b2amd_TapeBody& b = rb[s - 1];
		Body<T>& d = r.bodies[s];
		if ( initState != nullptr )
		{
			const double* x = initState + 7 * ( s - 1 );
			d.c = V2<T>{ FromD<T>( x[0] ), FromD<T>( x[1] ) };
			d.q = R2<T>{ FromD<T>( x[2] ), FromD<T>( x[3] ) };
			d.v = V2<T>{ FromD<T>( x[4] ), FromD<T>( x[5] ) };
			d.w = FromD<T>( x[6] );
		}
		else
		{
			d.c = Vk<T>( b.center );
			d.q = R2<T>{ K<T>( b.rotation.c ), K<T>( b.rotation.s ) };
			d.v = Vk<T>( b.linearVelocity );
			d.w = K<T>( b.angularVelocity );
		}
		d.dp = Vk<T>( { 0.0f, 0.0f } );
		d.dq = R2<T>{ K<T>( 1.0f ), K<T>( 0.0f ) };
		d.invMass = b.invMass;
		d.invI = b.invInertia;
		d.gravityScale = b.gravityScale;
		d.linearDamping = b.linearDamping;
		d.angularDamping = b.angularDamping;
		d.torque = b.torque;
		d.localCenter = b.localCenter;
		d.force = b.force;
		d.dynamic = ( b.flags & 0x00000200u ) != 0;
	}
}

// An independent variable: a tape node with no parents.
Rev RevVar( double v )
{
	g_tape->a.push_back( -1 );
	g_tape->b.push_back( -1 );
	g_tape->da.push_back( 0.0 );
	g_tape->db.push_back( 0.0 );
	return Rev{ v, (int)g_tape->a.size() - 1 };
}

} // namespace

extern "C" {

// Replay one recorded world.
//   mode 0: float replay; counts[0] = records it does not reproduce bit for bit
//   mode 1: float64 replay from initState (7 doubles per slot of record 0: center, rotation,
//           velocity, angular velocity) and shapeFriction
//   mode 2: mode 1 plus reverse-mode gradients of the loss into gradState / gradFriction
// loss = |center of targetSlot - goal|^2 after the last step. counts[1]: manifold point counts or
// feature ids that differ from the record; sig: decision signature of the run.
int b2o_adjoint_replay( int mode, int stepCount, const b2amd_TapeStep* headers, const b2amd_TapeBody* bodies, const int* bodyStart,
						const b2amd_TapeConstraint* cons, const int* consStart, const b2amd_Shape* shapes, const float* shapeFriction32,
						int shapeCount, const b2amd_TapeBody* statics, int staticCount, const double* initState,
						const double* shapeFriction, int targetSlot, const double* goal, double* loss, double* gradState,
						double* gradFriction, int* counts, uint64_t* sig )
{
	Input in{ stepCount, headers, bodies, bodyStart, cons, consStart, shapes, shapeFriction32, shapeCount, statics, staticCount };
	if ( mode == 0 )
	{
		Replay<float> r( in );
		InitBodies( r, nullptr );
		r.friction.resize( shapeCount );
		for ( int i = 0; i < shapeCount; ++i )
		{
			r.friction[i] = shapeFriction32[i];
		}
		for ( int k = 0; k < stepCount; ++k )
		{
			r.Step( k );
		}
		counts[0] = r.floatMismatch;
		counts[1] = r.structureMismatch;
		*sig = r.sig.h;
		const Body<float>& b = r.bodies[targetSlot];
		double dx = b.c.x - goal[0], dy = b.c.y - goal[1];
		*loss = dx * dx + dy * dy;
		return 0;
	}
	if ( mode == 1 )
	{
		Replay<double> r( in );
		InitBodies( r, initState );
		r.friction.assign( shapeFriction, shapeFriction + shapeCount );
		for ( int k = 0; k < stepCount; ++k )
		{
			r.Step( k );
		}
		counts[0] = 0;
		counts[1] = r.structureMismatch;
		*sig = r.sig.h;
		const Body<double>& b = r.bodies[targetSlot];
		double dx = b.c.x - goal[0], dy = b.c.y - goal[1];
		*loss = dx * dx + dy * dy;
		return 0;
	}

	RevTape tape;
	g_tape = &tape;
	Replay<Rev> r( in );
	int nb = headers[0].bodyCount;
	r.bodies.resize( nb + 1 );
	InitBodies( r, initState );
	// the initial state and the frictions are the independent variables (leaf nodes)
	std::vector<int> stateNodes( 7 * nb ), frictionNodes( shapeCount );
	for ( int s = 1; s <= nb; ++s )
	{
		Body<Rev>& b = r.bodies[s];
		Rev* vars[7] = { &b.c.x, &b.c.y, &b.q.c, &b.q.s, &b.v.x, &b.v.y, &b.w };
		for ( int k = 0; k < 7; ++k )
		{
			*vars[k] = RevVar( vars[k]->v );
			stateNodes[7 * ( s - 1 ) + k] = vars[k]->i;
		}
	}
	r.friction.resize( shapeCount );
	for ( int i = 0; i < shapeCount; ++i )
	{
		r.friction[i] = RevVar( shapeFriction[i] );
		frictionNodes[i] = r.friction[i].i;
	}
	// tape nodes of every body state at each step's start (and the final state)
	std::vector<int> stepNodes( (size_t)( stepCount + 1 ) * nb * 7, -1 );
	auto markStep = [&]( int k ) {
		for ( int s = 1; s <= nb; ++s )
		{
			Body<Rev>& bd = r.bodies[s];
			Rev* vars[7] = { &bd.c.x, &bd.c.y, &bd.q.c, &bd.q.s, &bd.v.x, &bd.v.y, &bd.w };
			for ( int c = 0; c < 7; ++c )
			{
				stepNodes[( (size_t)k * nb + ( s - 1 ) ) * 7 + c] = vars[c]->i;
			}
		}
	};
	for ( int k = 0; k < stepCount; ++k )
	{
		markStep( k );
		r.Step( k );
	}
	markStep( stepCount );
	const Body<Rev>& b = r.bodies[targetSlot];
	Rev dx = b.c.x - Rev{ goal[0], -1 };
	Rev dy = b.c.y - Rev{ goal[1], -1 };
	Rev L = dx * dx + dy * dy;
	*loss = L.v;
	counts[0] = 0;
	counts[1] = r.structureMismatch;
	*sig = r.sig.h;

	std::vector<double> adj( tape.a.size(), 0.0 );
	if ( L.i >= 0 )
	{
		adj[L.i] = 1.0;
	}
	for ( int n = (int)tape.a.size() - 1; n >= 0; --n )
	{
		double g = adj[n];
		if ( g == 0.0 )
		{
			continue;
		}
		if ( tape.a[n] >= 0 )
		{
			adj[tape.a[n]] += g * tape.da[n];
		}
		if ( tape.b[n] >= 0 )
		{
			adj[tape.b[n]] += g * tape.db[n];
		}
	}
	for ( int k = 0; k < 7 * nb; ++k )
	{
		gradState[k] = adj[stateNodes[k]];
	}
	for ( int i = 0; i < shapeCount; ++i )
	{
		gradFriction[i] = adj[frictionNodes[i]];
	}
	if ( g_stepGrads != nullptr )
	{
		for ( size_t k = 0; k < stepNodes.size(); ++k )
		{
			g_stepGrads[k] = stepNodes[k] >= 0 ? adj[stepNodes[k]] : 0.0;
		}
	}
	g_tape = nullptr;
	return 0;
}

// Branch pattern of the next replays: mode 0 off, 1 record (the pattern is kept), 2 force the kept
// pattern. b2o_adjoint_decisions_info reports the pattern length, forced branches the run's own
// values disagreed with, and whether a forced run asked for more branches than were kept.
static Decisions s_decisions;

void b2o_adjoint_decisions( int mode )
{
	s_decisions.mode = mode;
	s_decisions.pos = 0;
	s_decisions.flips = 0;
	s_decisions.overrun = false;
	if ( mode == 1 )
	{
		s_decisions.bits.clear();
	}
	g_dec = mode == 0 ? nullptr : &s_decisions;
}

void b2o_adjoint_decisions_info( int* length, int* flips, int* overrun, int* consumed )
{
	*length = (int)s_decisions.bits.size();
	*flips = s_decisions.flips;
	*overrun = s_decisions.overrun ? 1 : 0;
	*consumed = (int)s_decisions.pos;
}

// Gradients of the loss with respect to every body state at each step's start (steps + 1 blocks
// of bodies x 7, the last one the final state), from the next mode-2 replay.
void b2o_adjoint_capture_step_grads( double* out )
{
	g_stepGrads = out;
}

} // extern "C"
