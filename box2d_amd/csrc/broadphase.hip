// GPU broadphase pair query: a Morton-ordered linear BVH over every proxy of a world.
//
// The reference finds new pairs by querying its three dynamic trees with the fat AABB of every
// proxy in the move buffer (src/broad_phase.c:333-399, src/dynamic_tree.c:1085-1135) and filtering
// each hit in b2PairQueryCallback (src/broad_phase.c:172-324). Here the overlap search runs on the
// GPU: the fat AABBs of all proxies are sorted along a 2D Morton curve, a binary radix tree is
// built over the sorted codes (Karras, "Maximizing Parallelism in the Construction of BVHs,
// Octrees, and k-d Trees", HPG 2012), node boxes are refitted bottom-up, and one lane per moved
// proxy walks the tree. Each lane applies the callback's order-free rules (self, which trees a
// proxy type queries, the moved-proxy duplicate rule) and appends its hits. The candidate set is
// exactly the set of leaves the reference queries visit and keep at that point; the host sorts
// each query's candidates into the reference's visit order and applies the remaining filters.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "device.h"

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

constexpr int kLbvhBlock = 256;
constexpr int kLbvhStack = 96;

__device__ __forceinline__ bool Overlaps( const b2AABB& a, const b2AABB& b )
{
	// b2AABB_Overlaps (include/box2d/math_functions.h)
	return !( b.lowerBound.x > a.upperBound.x || b.lowerBound.y > a.upperBound.y || a.lowerBound.x > b.upperBound.x ||
			  a.lowerBound.y > b.upperBound.y );
}

// Scene bounds of the box centers (one workgroup).
__global__ __launch_bounds__( 1024 ) void LbvhBoundsKernel( const LbvhProxy* proxies, int n, float* bounds )
{
	__shared__ float s[4][1024];
	float lx = FLT_MAX, ly = FLT_MAX, ux = -FLT_MAX, uy = -FLT_MAX;
	for ( int i = threadIdx.x; i < n; i += blockDim.x )
	{
		b2AABB a = proxies[i].aabb;
		float cx = 0.5f * ( a.lowerBound.x + a.upperBound.x );
		float cy = 0.5f * ( a.lowerBound.y + a.upperBound.y );
		lx = fminf( lx, cx );
		ly = fminf( ly, cy );
		ux = fmaxf( ux, cx );
		uy = fmaxf( uy, cy );
	}
	s[0][threadIdx.x] = lx;
	s[1][threadIdx.x] = ly;
	s[2][threadIdx.x] = ux;
	s[3][threadIdx.x] = uy;
	__syncthreads();
	for ( int w = blockDim.x / 2; w > 0; w >>= 1 )
	{
		if ( threadIdx.x < w )
		{
			s[0][threadIdx.x] = fminf( s[0][threadIdx.x], s[0][threadIdx.x + w] );
			s[1][threadIdx.x] = fminf( s[1][threadIdx.x], s[1][threadIdx.x + w] );
			s[2][threadIdx.x] = fmaxf( s[2][threadIdx.x], s[2][threadIdx.x + w] );
			s[3][threadIdx.x] = fmaxf( s[3][threadIdx.x], s[3][threadIdx.x + w] );
		}
		__syncthreads();
	}
	if ( threadIdx.x < 4 )
	{
		bounds[threadIdx.x] = s[threadIdx.x][0];
	}
}

__device__ __forceinline__ uint32_t SpreadBits( uint32_t v )
{
	v &= 0xFFFF;
	v = ( v | ( v << 8 ) ) & 0x00FF00FF;
	v = ( v | ( v << 4 ) ) & 0x0F0F0F0F;
	v = ( v | ( v << 2 ) ) & 0x33333333;
	v = ( v | ( v << 1 ) ) & 0x55555555;
	return v;
}

// 64-bit keys: 32-bit Morton code of the box center above the proxy index (unique keys).
__global__ void LbvhMortonKernel( const LbvhProxy* proxies, int n, const float* bounds, uint64_t* keys )
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if ( i >= n )
	{
		return;
	}
	b2AABB a = proxies[i].aabb;
	float cx = 0.5f * ( a.lowerBound.x + a.upperBound.x );
	float cy = 0.5f * ( a.lowerBound.y + a.upperBound.y );
	float ex = bounds[2] - bounds[0];
	float ey = bounds[3] - bounds[1];
	float fx = ex > 0.0f ? ( cx - bounds[0] ) / ex : 0.0f;
	float fy = ey > 0.0f ? ( cy - bounds[1] ) / ey : 0.0f;
	uint32_t qx = (uint32_t)fminf( fmaxf( fx * 65536.0f, 0.0f ), 65535.0f );
	uint32_t qy = (uint32_t)fminf( fmaxf( fy * 65536.0f, 0.0f ), 65535.0f );
	uint32_t code = ( SpreadBits( qx ) << 1 ) | SpreadBits( qy );
	keys[i] = ( (uint64_t)code << 32 ) | (uint32_t)i;
}

struct LbvhTree
{
	b2AABB* nodeBox; // internal nodes, n - 1
	int* left;		 // child encoding: >= 0 internal node, < 0 leaf ~k
	int* right;
	int* parent;	   // internal parents (n - 1), then leaf parents (n)
	int* flags;		   // refit arrival counters, n - 1
	b2AABB* leafBox;   // sorted leaves
	LbvhProxy* leaves; // sorted proxies
	int n;
};

__global__ void LbvhLeafKernel( const LbvhProxy* proxies, const uint64_t* sortedKeys, LbvhTree t )
{
	int k = blockIdx.x * blockDim.x + threadIdx.x;
	if ( k < t.n )
	{
		int idx = (int)( sortedKeys[k] & 0xFFFFFFFFull );
		t.leaves[k] = proxies[idx];
		t.leafBox[k] = proxies[idx].aabb;
	}
	if ( k < t.n - 1 )
	{
		t.flags[k] = 0;
	}
}

__device__ __forceinline__ int Delta( const uint64_t* keys, int n, int i, int j )
{
	if ( j < 0 || j >= n )
	{
		return -1;
	}
	return __clzll( keys[i] ^ keys[j] );
}

// Binary radix tree over the sorted unique keys (Karras 2012, algorithm 1).
__global__ void LbvhBuildKernel( const uint64_t* keys, LbvhTree t )
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	int n = t.n;
	if ( i >= n - 1 )
	{
		return;
	}
	int d = Delta( keys, n, i, i + 1 ) - Delta( keys, n, i, i - 1 ) >= 0 ? 1 : -1;
	int dmin = Delta( keys, n, i, i - d );
	int lmax = 2;
	while ( Delta( keys, n, i, i + lmax * d ) > dmin )
	{
		lmax <<= 1;
	}
	int l = 0;
	for ( int step = lmax >> 1; step >= 1; step >>= 1 )
	{
		if ( Delta( keys, n, i, i + ( l + step ) * d ) > dmin )
		{
			l += step;
		}
	}
	int j = i + l * d;
	int dnode = Delta( keys, n, i, j );
	int s = 0;
	int step = l;
	do
	{
		step = ( step + 1 ) >> 1;
		if ( Delta( keys, n, i, i + ( s + step ) * d ) > dnode )
		{
			s += step;
		}
	}
	while ( step > 1 );
	int gamma = i + s * d + ( d < 0 ? d : 0 );
	int lo = i < j ? i : j;
	int hi = i < j ? j : i;
	int leftChild = lo == gamma ? ~gamma : gamma;
	int rightChild = hi == gamma + 1 ? ~( gamma + 1 ) : gamma + 1;
	t.left[i] = leftChild;
	t.right[i] = rightChild;
	if ( leftChild >= 0 )
	{
		t.parent[leftChild] = i;
	}
	else
	{
		t.parent[n - 1 + ~leftChild] = i;
	}
	if ( rightChild >= 0 )
	{
		t.parent[rightChild] = i;
	}
	else
	{
		t.parent[n - 1 + ~rightChild] = i;
	}
	if ( i == 0 )
	{
		t.parent[0] = -1;
	}
}

__device__ __forceinline__ float CoherentLoad( const float* p )
{
	return __uint_as_float( __hip_atomic_load( reinterpret_cast<const unsigned int*>( p ), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT ) );
}

__device__ __forceinline__ b2AABB ChildBox( const LbvhTree& t, int child )
{
	const float* p = child >= 0 ? &t.nodeBox[child].lowerBound.x : &t.leafBox[~child].lowerBound.x;
	b2AABB b;
	b.lowerBound.x = CoherentLoad( p + 0 );
	b.lowerBound.y = CoherentLoad( p + 1 );
	b.upperBound.x = CoherentLoad( p + 2 );
	b.upperBound.y = CoherentLoad( p + 3 );
	return b;
}

// Bottom-up refit: the second lane to reach a node unions its children and climbs on.
__global__ void LbvhRefitKernel( LbvhTree t )
{
	int k = blockIdx.x * blockDim.x + threadIdx.x;
	if ( k >= t.n || t.n < 2 )
	{
		return;
	}
	int node = t.parent[t.n - 1 + k];
	while ( node >= 0 )
	{
		__threadfence();
		int old = atomicAdd( t.flags + node, 1 );
		if ( old == 0 )
		{
			return;
		}
		b2AABB a = ChildBox( t, t.left[node] );
		b2AABB b = ChildBox( t, t.right[node] );
		b2AABB u;
		u.lowerBound.x = fminf( a.lowerBound.x, b.lowerBound.x );
		u.lowerBound.y = fminf( a.lowerBound.y, b.lowerBound.y );
		u.upperBound.x = fmaxf( a.upperBound.x, b.upperBound.x );
		u.upperBound.y = fmaxf( a.upperBound.y, b.upperBound.y );
		t.nodeBox[node] = u;
		node = t.parent[node];
	}
}

struct LbvhQueryArgs
{
	LbvhTree t;
	const LbvhProxy* queries;
	int queryCount;
	const uint8_t* moved; // by shape id: the shape's proxy is in the move buffer
	PairCandidate* out;
	int capacity;
	int* counters; // [0] candidates, [1] stack overflow
};

// Hit rules of b2PairQueryCallback that do not depend on host state (src/broad_phase.c:172-216).
__device__ __forceinline__ void LbvhVisit( const LbvhQueryArgs& a, int q, const LbvhProxy& query, const LbvhProxy& hit )
{
	int queryType = query.proxyKey & 3;
	int hitType = hit.proxyKey & 3;
	if ( hit.proxyKey == query.proxyKey )
	{
		return;
	}
	if ( queryType == b2_dynamicBody )
	{
		if ( hitType == b2_dynamicBody && hit.proxyKey < query.proxyKey && a.moved[hit.shapeId] )
		{
			return;
		}
	}
	else
	{
		// static and kinematic proxies only query the dynamic tree
		if ( hitType != b2_dynamicBody || a.moved[hit.shapeId] )
		{
			return;
		}
	}
	int slot = atomicAdd( a.counters, 1 );
	if ( slot < a.capacity )
	{
		a.out[slot] = PairCandidate{ q, hit.proxyKey, hit.shapeId };
	}
}

__global__ void LbvhQueryKernel( LbvhQueryArgs a )
{
	int q = blockIdx.x * blockDim.x + threadIdx.x;
	if ( q >= a.queryCount )
	{
		return;
	}
	const LbvhProxy query = a.queries[q];
	const b2AABB box = query.aabb;
	const LbvhTree& t = a.t;
	if ( t.n == 1 )
	{
		if ( Overlaps( t.leafBox[0], box ) )
		{
			LbvhVisit( a, q, query, t.leaves[0] );
		}
		return;
	}
	int stack[kLbvhStack];
	int sp = 0;
	stack[sp++] = 0;
	while ( sp > 0 )
	{
		int node = stack[--sp];
		int children[2] = { t.left[node], t.right[node] };
		for ( int c = 0; c < 2; ++c )
		{
			int child = children[c];
			if ( child < 0 )
			{
				if ( Overlaps( t.leafBox[~child], box ) )
				{
					LbvhVisit( a, q, query, t.leaves[~child] );
				}
			}
			else if ( Overlaps( t.nodeBox[child], box ) )
			{
				if ( sp < kLbvhStack )
				{
					stack[sp++] = child;
				}
				else
				{
					atomicAdd( a.counters + 1, 1 );
				}
			}
		}
	}
}

// ---------------------------------------------------------------------------------------------
// Host side

struct PairFinder::Impl
{
	hipStream_t stream = nullptr;
	size_t proxyCap = 0, queryCap = 0, candCap = 0, tempBytes = 0;
	LbvhProxy* proxies = nullptr;
	LbvhProxy* queries = nullptr;
	uint8_t* moved = nullptr;
	size_t movedCap = 0;
	uint64_t* keys = nullptr;
	uint64_t* sorted = nullptr;
	void* temp = nullptr;
	float* bounds = nullptr;
	LbvhTree tree = {};
	PairCandidate* cands = nullptr;
	int* counters = nullptr;
	int* hostCounters = nullptr;
	PairCandidate* hostCands = nullptr;
	size_t hostCandCap = 0;

	void Reserve( size_t n )
	{
		if ( n <= proxyCap )
		{
			return;
		}
		size_t cap = std::max<size_t>( 256, n + n / 2 );
		hipFree( proxies );
		hipFree( keys );
		hipFree( sorted );
		hipFree( tree.nodeBox );
		hipFree( tree.left );
		hipFree( tree.right );
		hipFree( tree.parent );
		hipFree( tree.flags );
		hipFree( tree.leafBox );
		hipFree( tree.leaves );
		HIP_CHECK( hipMalloc( &proxies, cap * sizeof( LbvhProxy ) ) );
		HIP_CHECK( hipMalloc( &keys, cap * sizeof( uint64_t ) ) );
		HIP_CHECK( hipMalloc( &sorted, cap * sizeof( uint64_t ) ) );
		HIP_CHECK( hipMalloc( &tree.nodeBox, cap * sizeof( b2AABB ) ) );
		HIP_CHECK( hipMalloc( &tree.left, cap * sizeof( int ) ) );
		HIP_CHECK( hipMalloc( &tree.right, cap * sizeof( int ) ) );
		HIP_CHECK( hipMalloc( &tree.parent, 2 * cap * sizeof( int ) ) );
		HIP_CHECK( hipMalloc( &tree.flags, cap * sizeof( int ) ) );
		HIP_CHECK( hipMalloc( &tree.leafBox, cap * sizeof( b2AABB ) ) );
		HIP_CHECK( hipMalloc( &tree.leaves, cap * sizeof( LbvhProxy ) ) );
		proxyCap = cap;
		size_t need = 0;
		HIP_CHECK( hipcub::DeviceRadixSort::SortKeys( nullptr, need, keys, sorted, (int)cap, 0, 64, stream ) );
		if ( need > tempBytes )
		{
			hipFree( temp );
			HIP_CHECK( hipMalloc( &temp, need ) );
			tempBytes = need;
		}
	}
};

PairFinder::PairFinder()
{
	impl = new Impl();
	HIP_CHECK( hipStreamCreate( &impl->stream ) );
	HIP_CHECK( hipMalloc( &impl->bounds, 4 * sizeof( float ) ) );
	HIP_CHECK( hipMalloc( &impl->counters, 2 * sizeof( int ) ) );
	HIP_CHECK( hipHostMalloc( &impl->hostCounters, 2 * sizeof( int ), hipHostMallocDefault ) );
}

PairFinder::~PairFinder()
{
	Impl* m = impl;
	hipFree( m->proxies );
	hipFree( m->queries );
	hipFree( m->moved );
	hipFree( m->keys );
	hipFree( m->sorted );
	hipFree( m->temp );
	hipFree( m->bounds );
	hipFree( m->tree.nodeBox );
	hipFree( m->tree.left );
	hipFree( m->tree.right );
	hipFree( m->tree.parent );
	hipFree( m->tree.flags );
	hipFree( m->tree.leafBox );
	hipFree( m->tree.leaves );
	hipFree( m->cands );
	hipFree( m->counters );
	hipHostFree( m->hostCounters );
	hipHostFree( m->hostCands );
	hipStreamDestroy( m->stream );
	delete m;
}

void PairFinder::Find( const std::vector<LbvhProxy>& proxies, const std::vector<LbvhProxy>& queries,
					   const std::vector<uint8_t>& moved, std::vector<PairCandidate>* out )
{
	Impl* m = impl;
	hipStream_t s = m->stream;
	out->clear();
	int n = (int)proxies.size();
	int nq = (int)queries.size();
	if ( n == 0 || nq == 0 )
	{
		return;
	}
	m->Reserve( n );
	if ( (size_t)nq > m->queryCap )
	{
		hipFree( m->queries );
		m->queryCap = std::max<size_t>( 256, nq + nq / 2 );
		HIP_CHECK( hipMalloc( &m->queries, m->queryCap * sizeof( LbvhProxy ) ) );
	}
	if ( moved.size() > m->movedCap )
	{
		hipFree( m->moved );
		m->movedCap = std::max<size_t>( 1024, moved.size() + moved.size() / 2 );
		HIP_CHECK( hipMalloc( &m->moved, m->movedCap ) );
	}
	if ( m->candCap == 0 )
	{
		m->candCap = std::max<size_t>( 4096, 8 * (size_t)nq );
		HIP_CHECK( hipMalloc( &m->cands, m->candCap * sizeof( PairCandidate ) ) );
	}

	HIP_CHECK( hipMemcpyAsync( m->proxies, proxies.data(), n * sizeof( LbvhProxy ), hipMemcpyHostToDevice, s ) );
	HIP_CHECK( hipMemcpyAsync( m->queries, queries.data(), nq * sizeof( LbvhProxy ), hipMemcpyHostToDevice, s ) );
	HIP_CHECK( hipMemcpyAsync( m->moved, moved.data(), moved.size(), hipMemcpyHostToDevice, s ) );

	LbvhTree t = m->tree;
	t.n = n;
	LbvhBoundsKernel<<<1, 1024, 0, s>>>( m->proxies, n, m->bounds );
	int blocks = ( n + kLbvhBlock - 1 ) / kLbvhBlock;
	LbvhMortonKernel<<<blocks, kLbvhBlock, 0, s>>>( m->proxies, n, m->bounds, m->keys );
	size_t temp = m->tempBytes;
	HIP_CHECK( hipcub::DeviceRadixSort::SortKeys( m->temp, temp, m->keys, m->sorted, n, 0, 64, s ) );
	LbvhLeafKernel<<<blocks, kLbvhBlock, 0, s>>>( m->proxies, m->sorted, t );
	if ( n > 1 )
	{
		LbvhBuildKernel<<<blocks, kLbvhBlock, 0, s>>>( m->sorted, t );
		LbvhRefitKernel<<<blocks, kLbvhBlock, 0, s>>>( t );
	}
	HIP_CHECK( hipGetLastError() );

	for ( ;; )
	{
		LbvhQueryArgs qa;
		qa.t = t;
		qa.queries = m->queries;
		qa.queryCount = nq;
		qa.moved = m->moved;
		qa.out = m->cands;
		qa.capacity = (int)m->candCap;
		qa.counters = m->counters;
		HIP_CHECK( hipMemsetAsync( m->counters, 0, 2 * sizeof( int ), s ) );
		LbvhQueryKernel<<<( nq + kLbvhBlock - 1 ) / kLbvhBlock, kLbvhBlock, 0, s>>>( qa );
		HIP_CHECK( hipGetLastError() );
		HIP_CHECK( hipMemcpyAsync( m->hostCounters, m->counters, 2 * sizeof( int ), hipMemcpyDeviceToHost, s ) );
		HIP_CHECK( hipStreamSynchronize( s ) );
		int count = m->hostCounters[0];
		if ( m->hostCounters[1] != 0 )
		{
			fprintf( stderr, "box2d_amd: LBVH query stack overflow\n" );
			abort();
		}
		if ( count > (int)m->candCap )
		{
			hipFree( m->cands );
			m->candCap = (size_t)count + count / 2;
			HIP_CHECK( hipMalloc( &m->cands, m->candCap * sizeof( PairCandidate ) ) );
			continue;
		}
		if ( (size_t)count > m->hostCandCap )
		{
			hipHostFree( m->hostCands );
			m->hostCandCap = (size_t)count + count / 2;
			HIP_CHECK( hipHostMalloc( &m->hostCands, m->hostCandCap * sizeof( PairCandidate ), hipHostMallocDefault ) );
		}
		if ( count > 0 )
		{
			HIP_CHECK( hipMemcpyAsync( m->hostCands, m->cands, count * sizeof( PairCandidate ), hipMemcpyDeviceToHost, s ) );
			HIP_CHECK( hipStreamSynchronize( s ) );
		}
		out->assign( m->hostCands, m->hostCands + count );
		break;
	}
}

} // namespace b2g
