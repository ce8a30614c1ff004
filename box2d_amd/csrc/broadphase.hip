// GPU broadphase pair query: a Morton-ordered linear BVH over every proxy of a world.
//
// The reference finds new pairs by querying its three dynamic trees with the fat AABB of every
// proxy in the move buffer (src/broad_phase.c:333-399, src/dynamic_tree.c:1085-1135) and filtering
// each hit in b2PairQueryCallback (src/broad_phase.c:172-324). Here the overlap search runs on the
// GPU: the fat AABBs of all proxies are sorted along a 2D Morton curve and a binary radix tree is
// built over the sorted codes (Karras, "Maximizing Parallelism in the Construction of BVHs,
// Octrees, and k-d Trees", HPG 2012). Each internal node's box is the union of its sorted-leaf
// range, taken from leaf boxes and 64-leaf chunk boxes by one wavefront per node (exact: min and
// max are order-free), so no bottom-up refit pass is needed. Every node carries a rope, the next
// node of a depth-first walk once its subtree is skipped, and one lane per moved proxy walks the
// tree without a stack. Each lane applies the callback's order-free rules (self, which trees a
// proxy type queries, the moved-proxy duplicate rule) and emits (query, proxy index) keys, which
// are radix sorted on the device. The caller numbers the proxies in the reference's visit order
// (kinematic tree, static tree, dynamic tree, each depth first), so the sorted keys list every
// query's candidates in the order the reference's queries visit them; the host applies the
// remaining filters in that order.
//
// A query box much larger than the typical proxy (a rotating container wall, a spinner arm)
// overlaps most of the tree, and one lane walking thousands of nodes would stall the whole query.
// Queries whose box perimeter exceeds kLargeQueryFactor times the mean proxy perimeter are
// therefore not walked: one lane per proxy tests every such query instead. Both paths apply the
// same rules and emit the same keys, so the sorted result does not depend on the split.

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
constexpr int kLbvhChunk = 64; // leaves per chunk box (one wavefront)
constexpr float kLargeQueryFactor = 8.0f;

__device__ __forceinline__ float Perimeter( const b2AABB& a )
{
	return ( a.upperBound.x - a.lowerBound.x ) + ( a.upperBound.y - a.lowerBound.y );
}

__device__ __forceinline__ bool Overlaps( const b2AABB& a, const b2AABB& b )
{
	// b2AABB_Overlaps (include/box2d/math_functions.h)
	return !( b.lowerBound.x > a.upperBound.x || b.lowerBound.y > a.upperBound.y || a.lowerBound.x > b.upperBound.x ||
			  a.lowerBound.y > b.upperBound.y );
}

__device__ __forceinline__ b2AABB EmptyBox()
{
	b2AABB b;
	b.lowerBound = { FLT_MAX, FLT_MAX };
	b.upperBound = { -FLT_MAX, -FLT_MAX };
	return b;
}

__device__ __forceinline__ void Grow( b2AABB& a, const b2AABB& b )
{
	a.lowerBound.x = fminf( a.lowerBound.x, b.lowerBound.x );
	a.lowerBound.y = fminf( a.lowerBound.y, b.lowerBound.y );
	a.upperBound.x = fmaxf( a.upperBound.x, b.upperBound.x );
	a.upperBound.y = fmaxf( a.upperBound.y, b.upperBound.y );
}

// Union over the 64 lanes of a wavefront (every lane must be active).
__device__ __forceinline__ b2AABB WaveUnion( b2AABB a )
{
	for ( int m = 1; m < 64; m <<= 1 )
	{
		b2AABB o;
		o.lowerBound.x = __shfl_xor( a.lowerBound.x, m, 64 );
		o.lowerBound.y = __shfl_xor( a.lowerBound.y, m, 64 );
		o.upperBound.x = __shfl_xor( a.upperBound.x, m, 64 );
		o.upperBound.y = __shfl_xor( a.upperBound.y, m, 64 );
		Grow( a, o );
	}
	return a;
}

// Scene bounds of the box centers and the large-query perimeter threshold (one workgroup):
// bounds[0..3] = center bounds, bounds[4] = kLargeQueryFactor x mean proxy perimeter.
__global__ __launch_bounds__( 1024 ) void LbvhBoundsKernel( const LbvhProxy* proxies, int n, float* bounds )
{
	__shared__ float s[5][1024];
	float lx = FLT_MAX, ly = FLT_MAX, ux = -FLT_MAX, uy = -FLT_MAX, per = 0.0f;
	for ( int i = threadIdx.x; i < n; i += blockDim.x )
	{
		b2AABB a = proxies[i].aabb;
		float cx = 0.5f * ( a.lowerBound.x + a.upperBound.x );
		float cy = 0.5f * ( a.lowerBound.y + a.upperBound.y );
		lx = fminf( lx, cx );
		ly = fminf( ly, cy );
		ux = fmaxf( ux, cx );
		uy = fmaxf( uy, cy );
		per += Perimeter( a );
	}
	s[0][threadIdx.x] = lx;
	s[1][threadIdx.x] = ly;
	s[2][threadIdx.x] = ux;
	s[3][threadIdx.x] = uy;
	s[4][threadIdx.x] = per;
	__syncthreads();
	for ( int w = blockDim.x / 2; w > 0; w >>= 1 )
	{
		if ( threadIdx.x < w )
		{
			s[0][threadIdx.x] = fminf( s[0][threadIdx.x], s[0][threadIdx.x + w] );
			s[1][threadIdx.x] = fminf( s[1][threadIdx.x], s[1][threadIdx.x + w] );
			s[2][threadIdx.x] = fmaxf( s[2][threadIdx.x], s[2][threadIdx.x + w] );
			s[3][threadIdx.x] = fmaxf( s[3][threadIdx.x], s[3][threadIdx.x + w] );
			s[4][threadIdx.x] = s[4][threadIdx.x] + s[4][threadIdx.x + w];
		}
		__syncthreads();
	}
	if ( threadIdx.x < 4 )
	{
		bounds[threadIdx.x] = s[threadIdx.x][0];
	}
	if ( threadIdx.x == 4 )
	{
		bounds[4] = kLargeQueryFactor * s[4][0] / (float)n;
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

// 32-bit Morton code of each box center; the proxy index rides along as the sort value (the sort
// is stable, so equal codes stay in proxy order).
__global__ void LbvhMortonKernel( const LbvhProxy* proxies, int n, const float* bounds, uint32_t* codes, uint32_t* index )
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
	codes[i] = ( SpreadBits( qx ) << 1 ) | SpreadBits( qy );
	index[i] = (uint32_t)i;
}

// Tree nodes, 32 bytes: internal nodes 0..n-2 (node 0 is the root), leaf k of the sorted order at
// n-1+k (for n == 1 the single leaf is node 0).
struct LbvhNode
{
	b2AABB box;
	int child;	  // internal: node index of the left child; leaf: ~(proxy index)
	int rope;	  // next node once this subtree is skipped, -1 at the end of the walk
	int proxyKey; // leaves only
	int shapeId;  // leaves only
};

struct LbvhTree
{
	LbvhNode* nodes;	 // 2n - 1
	b2AABB* chunkBox;	 // union of each 64 consecutive sorted leaves
	int* rightOf;		 // n - 1: node index of the right child of the internal node that splits at k | k+1
	int* rangeLo;		 // n - 1: sorted-leaf range of each internal node
	int* rangeHi;
	const uint32_t* codes; // sorted Morton codes
	int n;
};

// Sorted leaves and their chunk boxes (blocks of whole wavefronts, so each wavefront is a chunk).
__global__ __launch_bounds__( kLbvhBlock ) void LbvhLeafKernel( const LbvhProxy* proxies, const uint32_t* sortedIndex, LbvhTree t )
{
	int k = blockIdx.x * blockDim.x + threadIdx.x;
	b2AABB box = EmptyBox();
	if ( k < t.n )
	{
		int idx = (int)sortedIndex[k];
		const LbvhProxy p = proxies[idx];
		box = p.aabb;
		LbvhNode& node = t.nodes[t.n - 1 + k];
		node.box = box;
		node.child = ~idx;
		node.proxyKey = p.proxyKey;
		node.shapeId = p.shapeId;
	}
	box = WaveUnion( box );
	if ( ( threadIdx.x & 63 ) == 0 && k < t.n )
	{
		t.chunkBox[k / kLbvhChunk] = box;
	}
}

// Common prefix length of sorted keys i and j, -1 outside [0, n); keys are (code, position) so
// they are unique.
__device__ __forceinline__ int Delta( const uint32_t* codes, int n, int i, int j )
{
	if ( j < 0 || j >= n )
	{
		return -1;
	}
	uint32_t a = codes[i], b = codes[j];
	if ( a != b )
	{
		return __clz( a ^ b );
	}
	return 32 + __clz( (uint32_t)i ^ (uint32_t)j );
}

// Binary radix tree over the sorted keys (Karras 2012, algorithm 1).
__global__ void LbvhBuildKernel( LbvhTree t )
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	int n = t.n;
	if ( i >= n - 1 )
	{
		return;
	}
	const uint32_t* codes = t.codes;
	int d = Delta( codes, n, i, i + 1 ) - Delta( codes, n, i, i - 1 ) >= 0 ? 1 : -1;
	int dmin = Delta( codes, n, i, i - d );
	int lmax = 2;
	while ( Delta( codes, n, i, i + lmax * d ) > dmin )
	{
		lmax <<= 1;
	}
	int l = 0;
	for ( int step = lmax >> 1; step >= 1; step >>= 1 )
	{
		if ( Delta( codes, n, i, i + ( l + step ) * d ) > dmin )
		{
			l += step;
		}
	}
	int j = i + l * d;
	int dnode = Delta( codes, n, i, j );
	int s = 0;
	int step = l;
	do
	{
		step = ( step + 1 ) >> 1;
		if ( Delta( codes, n, i, i + ( s + step ) * d ) > dnode )
		{
			s += step;
		}
	}
	while ( step > 1 );
	int gamma = i + s * d + ( d < 0 ? d : 0 );
	int lo = i < j ? i : j;
	int hi = i < j ? j : i;
	int left = lo == gamma ? n - 1 + gamma : gamma;
	int right = hi == gamma + 1 ? n - 1 + gamma + 1 : gamma + 1;
	t.nodes[i].child = left;
	t.nodes[i].proxyKey = -1;
	t.nodes[i].shapeId = -1;
	t.rightOf[gamma] = right;
	t.rangeLo[i] = lo;
	t.rangeHi[i] = hi;
}

// Ropes of every node and boxes of the internal nodes. Blocks [0, leafBlocks) take one leaf per
// lane; the rest take one internal node per wavefront.
__global__ __launch_bounds__( kLbvhBlock ) void LbvhFinishKernel( LbvhTree t, int leafBlocks )
{
	int n = t.n;
	if ( (int)blockIdx.x < leafBlocks )
	{
		int k = blockIdx.x * blockDim.x + threadIdx.x;
		if ( k < n )
		{
			t.nodes[n - 1 + k].rope = k < n - 1 ? t.rightOf[k] : -1;
		}
		return;
	}
	int i = ( blockIdx.x - leafBlocks ) * ( blockDim.x / 64 ) + threadIdx.x / 64;
	int lane = threadIdx.x & 63;
	if ( i >= n - 1 )
	{
		return; // whole wavefront
	}
	int lo = t.rangeLo[i], hi = t.rangeHi[i];
	b2AABB box = EmptyBox();
	int c0 = ( lo + kLbvhChunk - 1 ) / kLbvhChunk; // first whole chunk
	int c1 = ( hi + 1 ) / kLbvhChunk;			   // one past the last whole chunk
	if ( c0 >= c1 )
	{
		for ( int k = lo + lane; k <= hi; k += 64 )
		{
			Grow( box, t.nodes[n - 1 + k].box );
		}
	}
	else
	{
		for ( int k = lo + lane; k < c0 * kLbvhChunk; k += 64 )
		{
			Grow( box, t.nodes[n - 1 + k].box );
		}
		for ( int c = c0 + lane; c < c1; c += 64 )
		{
			Grow( box, t.chunkBox[c] );
		}
		for ( int k = c1 * kLbvhChunk + lane; k <= hi; k += 64 )
		{
			Grow( box, t.nodes[n - 1 + k].box );
		}
	}
	box = WaveUnion( box );
	if ( lane == 0 )
	{
		t.nodes[i].box = box;
		t.nodes[i].rope = hi < n - 1 ? t.rightOf[hi] : -1;
	}
}

struct LbvhQueryArgs
{
	const LbvhNode* nodes;
	int nodeCount;
	const LbvhProxy* queries;
	int queryCount;
	const uint8_t* moved; // by shape id: the shape's proxy is in the move buffer
	uint64_t* out;		  // (query << 32 | proxy index)
	int capacity;
	int* counters; // [0] candidates, [1] walks cut off by the step guard, [2] large queries
	const float* bounds; // [4]: large-query perimeter threshold
	uint8_t* large;		 // per query: 1 if tested by LbvhLargeQueryKernel instead of walked
	int* largeList;		 // the large queries' indices (counters[2] of them)
	const LbvhProxy* proxies;
	int proxyCount;
	const uint64_t* pairSlots; // mirror of the existing-pair set (nullptr: empty)
	uint64_t pairMask;
	const PairShapeFilter* filters; // by shape id (nullptr: not synced)
	const int* queryCountOnDevice;	// when set, the number of queries (queryCount is then a bound)
};

// Whether the shape pair already has a contact: the host set's probe sequence on its mirror.
__device__ __forceinline__ bool InPairSet( const LbvhQueryArgs& a, int shapeA, int shapeB )
{
	if ( a.pairSlots == nullptr )
	{
		return false;
	}
	uint64_t key = shapeA < shapeB ? ( (uint64_t)shapeA << 32 ) | (uint32_t)shapeB : ( (uint64_t)shapeB << 32 ) | (uint32_t)shapeA;
	uint64_t i = PairKeyHash( key ) & a.pairMask;
	for ( uint64_t probes = 0; probes <= a.pairMask; ++probes )
	{
		uint64_t k = a.pairSlots[i];
		if ( k == key )
		{
			return true;
		}
		if ( k == 0 )
		{
			return false;
		}
		i = ( i + 1 ) & a.pairMask;
	}
	return false;
}

// Appends a hit's key with one counter atomic per wavefront for all of its lanes' hits.
__device__ __forceinline__ void EmitHits( const LbvhQueryArgs& a, bool hit, uint64_t key )
{
	const int lane = threadIdx.x & 63;
	uint64_t mask = __ballot( hit );
	if ( mask != 0 )
	{
		int leader = __ffsll( (unsigned long long)mask ) - 1;
		int base = 0;
		if ( lane == leader )
		{
			base = atomicAdd( a.counters, __popcll( mask ) );
		}
		base = __shfl( base, leader, 64 );
		if ( hit )
		{
			int slot = base + __popcll( mask & ( ( 1ull << lane ) - 1ull ) );
			if ( slot < a.capacity )
			{
				a.out[slot] = key;
			}
		}
	}
}

// Hit rules of b2PairQueryCallback that do not depend on host state (src/broad_phase.c:172-216),
// and its existing-pair test (:218-226) on the mirrored pair set.
__device__ __forceinline__ bool LbvhKeep( const LbvhQueryArgs& a, int queryKey, int queryShape, int hitKey, int hitShape )
{
	int queryType = queryKey & 3;
	int hitType = hitKey & 3;
	if ( hitKey == queryKey )
	{
		return false;
	}
	if ( queryType == b2_dynamicBody )
	{
		if ( hitType == b2_dynamicBody && hitKey < queryKey && a.moved[hitShape] )
		{
			return false;
		}
	}
	else
	{
		// static and kinematic proxies only query the dynamic tree
		if ( hitType != b2_dynamicBody || a.moved[hitShape] )
		{
			return false;
		}
	}
	if ( a.filters != nullptr )
	{
		// same body and b2ShouldShapesCollide (src/broad_phase.c:245-264, src/shape.h:122-130)
		const PairShapeFilter fa = a.filters[queryShape];
		const PairShapeFilter fb = a.filters[hitShape];
		if ( fa.bodyId == fb.bodyId )
		{
			return false;
		}
		if ( fa.groupIndex == fb.groupIndex && fa.groupIndex != 0 )
		{
			if ( fa.groupIndex <= 0 )
			{
				return false;
			}
		}
		else if ( ( fa.maskBits & fb.categoryBits ) == 0 || ( fa.categoryBits & fb.maskBits ) == 0 )
		{
			return false;
		}
	}
	return InPairSet( a, queryShape, hitShape ) == false;
}

// One lane per query: stackless depth-first walk along the ropes. Every node is entered at most
// once, so the walk takes at most nodeCount steps; the guard only bounds a malformed tree.
__global__ __launch_bounds__( kLbvhBlock ) void LbvhQueryKernel( LbvhQueryArgs a )
{
	int q = blockIdx.x * blockDim.x + threadIdx.x;
	if ( q >= a.queryCount )
	{
		return;
	}
	if ( a.large[q] )
	{
		return;
	}
	const LbvhProxy query = a.queries[q];
	const b2AABB box = query.aabb;
	int node = 0;
	int steps = 0;
	while ( node >= 0 )
	{
		if ( ++steps > a.nodeCount )
		{
			atomicAdd( a.counters + 1, 1 );
			break;
		}
		const LbvhNode e = a.nodes[node];
		bool hit = false;
		int next = e.rope;
		if ( Overlaps( e.box, box ) )
		{
			if ( e.child < 0 )
			{
				hit = LbvhKeep( a, query.proxyKey, query.shapeId, e.proxyKey, e.shapeId );
			}
			else
			{
				next = e.child;
			}
		}
		EmitHits( a, hit, ( (uint64_t)q << 32 ) | (uint32_t)( ~e.child ) );
		node = next;
	}
}

// Marks and lists the queries whose box is too large to walk.
__global__ __launch_bounds__( kLbvhBlock ) void LbvhLargeListKernel( LbvhQueryArgs a )
{
	int q = blockIdx.x * blockDim.x + threadIdx.x;
	if ( q >= a.queryCount )
	{
		return;
	}
	bool large = Perimeter( a.queries[q].aabb ) > a.bounds[4];
	a.large[q] = large ? 1 : 0;
	if ( large )
	{
		a.largeList[atomicAdd( a.counters + 2, 1 )] = q;
	}
}

// Large queries, one lane per proxy: every proxy is tested against every large query box, with
// the walk's rules and keys.
__global__ __launch_bounds__( kLbvhBlock ) void LbvhLargeQueryKernel( LbvhQueryArgs a )
{
	__shared__ LbvhProxy sq[kLbvhBlock];
	__shared__ int sIndex[kLbvhBlock];
	int p = blockIdx.x * blockDim.x + threadIdx.x;
	int largeCount = a.counters[2];
	LbvhProxy proxy = {};
	if ( p < a.proxyCount )
	{
		proxy = a.proxies[p];
	}
	for ( int base = 0; base < largeCount; base += kLbvhBlock )
	{
		int m = min( kLbvhBlock, largeCount - base );
		__syncthreads();
		if ( (int)threadIdx.x < m )
		{
			int q = a.largeList[base + threadIdx.x];
			sIndex[threadIdx.x] = q;
			sq[threadIdx.x] = a.queries[q];
		}
		__syncthreads();
		for ( int j = 0; j < m; ++j )
		{
			bool hit = p < a.proxyCount && Overlaps( proxy.aabb, sq[j].aabb ) &&
					   LbvhKeep( a, sq[j].proxyKey, sq[j].shapeId, proxy.proxyKey, proxy.shapeId );
			EmitHits( a, hit, ( (uint64_t)sIndex[j] << 32 ) | (uint32_t)p );
		}
	}
}

// ---------------------------------------------------------------------------------------------
// Chunk scan: the default query.
//
// The staged proxies are in the reference trees' depth-first leaf order, which keeps neighbours
// in space near each other in the array, so 64-proxy chunks of it have tight boxes. One wavefront
// per chunk takes the union of its 64 boxes; then one wavefront per query tests 64 chunk boxes at
// a time against the query box and, for each overlapping chunk in ascending order, its 64
// proxies, one per lane. A query is a handful of wavefront-wide loads instead of one lane's
// dependent walk, and its hits come out in ascending proxy order (chunks ascending, lanes
// ascending, and each wavefront's appends at increasing counter values), so no sort is needed.

__global__ __launch_bounds__( kLbvhBlock ) void PairChunkBoxKernel( const LbvhProxy* proxies, int n, b2AABB* chunkBox )
{
	int chunk = ( blockIdx.x * blockDim.x + threadIdx.x ) >> 6;
	int lane = threadIdx.x & 63;
	if ( chunk * kLbvhChunk >= n )
	{
		return; // whole wavefront
	}
	int p = chunk * kLbvhChunk + lane;
	b2AABB box = p < n ? proxies[p].aabb : EmptyBox();
	box = WaveUnion( box );
	if ( lane == 0 )
	{
		chunkBox[chunk] = box;
	}
}

__global__ __launch_bounds__( kLbvhBlock ) void PairChunkQueryKernel( LbvhQueryArgs a, const b2AABB* chunkBox, int chunkCount )
{
	int q = ( blockIdx.x * blockDim.x + threadIdx.x ) >> 6;
	int lane = threadIdx.x & 63;
	int queryCount = a.queryCountOnDevice != nullptr ? min( *a.queryCountOnDevice, a.queryCount ) : a.queryCount;
	if ( q >= queryCount )
	{
		return; // whole wavefront
	}
	const LbvhProxy query = a.queries[q];
	const b2AABB box = query.aabb;
	for ( int base = 0; base < chunkCount; base += 64 )
	{
		int c = base + lane;
		bool overlap = c < chunkCount && Overlaps( chunkBox[c], box );
		uint64_t chunks = __ballot( overlap );
		while ( chunks != 0 )
		{
			int first = __ffsll( (unsigned long long)chunks ) - 1;
			chunks &= chunks - 1;
			int p = ( base + first ) * kLbvhChunk + lane;
			bool hit = false;
			if ( p < a.proxyCount )
			{
				const LbvhProxy proxy = a.proxies[p];
				hit = Overlaps( proxy.aabb, box ) && LbvhKeep( a, query.proxyKey, query.shapeId, proxy.proxyKey, proxy.shapeId );
			}
			EmitHits( a, hit, ( (uint64_t)q << 32 ) | (uint32_t)p );
		}
	}
}

// Queries of a speculative launch from a solve's enlarged shapes: each enlarged shape becomes a
// query with its new fat box, its staged proxy takes that box, and it is marked moved. shapeSlots
// maps a shape id to (proxy key, staged position).
__global__ __launch_bounds__( kLbvhBlock ) void PairSpeculativeBuildKernel( const int* enlargedShapes, const b2AABB* enlargedAABBs,
																		   const int* enlargedCount, int capacity, const ShapeSlot* shapeSlots,
																		   int shapeCount, int proxyCount, LbvhProxy* queries, LbvhProxy* proxies,
																		   uint8_t* moved, int* queryCount, int* error )
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	int count = *enlargedCount;
	if ( i == 0 )
	{
		*queryCount = count;
	}
	if ( i >= count || i >= capacity )
	{
		return;
	}
	int shapeId = enlargedShapes[i];
	b2AABB box = enlargedAABBs[i];
	ShapeSlot slot = shapeId >= 0 && shapeId < shapeCount ? shapeSlots[shapeId] : ShapeSlot{ -1, -1 };
	if ( slot.position < 0 || slot.position >= proxyCount )
	{
		// not a staged proxy: the result is incomplete and the caller queries again
		atomicAdd( error, 1 );
		queries[i] = LbvhProxy{ -1, -1, EmptyBox() };
		return;
	}
	queries[i] = LbvhProxy{ slot.proxyKey, shapeId, box };
	proxies[slot.position].aabb = box;
	moved[shapeId] = 1;
}

// ---------------------------------------------------------------------------------------------
// Host side

// Keys copied back with the candidate count, before the count is known (one round trip when the
// candidates fit).
constexpr int kEagerKeys = 4096;

struct PairFinder::Impl
{
	hipStream_t stream = nullptr;
	int mode = 0; // 0 chunk scan, 1 LBVH walk
	b2AABB* chunkBoxes = nullptr;
	size_t chunkCap = 0;
	ShapeSlot* shapeSlots = nullptr; // speculative launches: shape id -> (proxy key, staged position)
	size_t shapeSlotCap = 0;
	ShapeSlot* hostShapeSlots = nullptr;
	size_t hostShapeSlotCap = 0;
	bool speculative = false; // the launched query is a speculative one
	size_t proxyCap = 0, queryCap = 0, candCap = 0;
	LbvhProxy* proxies = nullptr;
	LbvhProxy* queries = nullptr;
	uint8_t* large = nullptr; // per query
	int* largeList = nullptr;
	uint8_t* moved = nullptr;
	size_t movedCap = 0;
	uint32_t* codes = nullptr;
	uint32_t* sortedCodes = nullptr;
	uint32_t* index = nullptr;
	uint32_t* sortedIndex = nullptr;
	void* temp = nullptr;
	size_t tempBytes = 0;
	float* bounds = nullptr;
	LbvhTree tree = {};
	uint64_t* cands = nullptr;
	uint64_t* sortedCands = nullptr;
	uint64_t* pairSlots = nullptr; // mirror of the existing-pair set
	size_t pairCap = 0;			   // allocated slots
	size_t pairSize = 0;		   // slots of the mirrored table (0: empty set)
	uint64_t* pairUpdates = nullptr; // (slot, key) pairs of a partial update
	size_t pairUpdateCap = 0;
	int* counters = nullptr;
	int* hostCounters = nullptr;

	// pinned host staging
	LbvhProxy* hostProxies = nullptr;
	size_t hostProxyCap = 0;
	LbvhProxy* hostQueries = nullptr;
	size_t hostQueryCap = 0;
	uint8_t* hostMoved = nullptr;
	size_t hostMovedCap = 0;
	uint64_t* hostCands = nullptr;
	size_t hostCandCap = 0;
	uint64_t* hostPairs = nullptr; // whole table or (slot, key) pairs
	size_t hostPairCap = 0;
	PairShapeFilter* filters = nullptr;
	size_t filterCap = 0;
	int filterCount = 0;
	PairShapeFilter* hostFilters = nullptr;
	size_t hostFilterCap = 0;

	// the launched query, finished by Finish
	int launchedProxies = 0;
	int launchedQueries = 0;
	void Query();

	void Temp( size_t need )
	{
		if ( need > tempBytes )
		{
			hipFree( temp );
			tempBytes = need + need / 2;
			HIP_CHECK( hipMalloc( &temp, tempBytes ) );
		}
	}

	void Reserve( size_t n )
	{
		if ( n <= proxyCap )
		{
			return;
		}
		size_t cap = std::max<size_t>( 256, n + n / 2 );
		hipFree( proxies );
		hipFree( codes );
		hipFree( sortedCodes );
		hipFree( index );
		hipFree( sortedIndex );
		hipFree( tree.nodes );
		hipFree( tree.chunkBox );
		hipFree( tree.rightOf );
		hipFree( tree.rangeLo );
		hipFree( tree.rangeHi );
		HIP_CHECK( hipMalloc( &proxies, cap * sizeof( LbvhProxy ) ) );
		HIP_CHECK( hipMalloc( &codes, cap * sizeof( uint32_t ) ) );
		HIP_CHECK( hipMalloc( &sortedCodes, cap * sizeof( uint32_t ) ) );
		HIP_CHECK( hipMalloc( &index, cap * sizeof( uint32_t ) ) );
		HIP_CHECK( hipMalloc( &sortedIndex, cap * sizeof( uint32_t ) ) );
		HIP_CHECK( hipMalloc( &tree.nodes, 2 * cap * sizeof( LbvhNode ) ) );
		HIP_CHECK( hipMalloc( &tree.chunkBox, ( cap / kLbvhChunk + 1 ) * sizeof( b2AABB ) ) );
		HIP_CHECK( hipMalloc( &tree.rightOf, cap * sizeof( int ) ) );
		HIP_CHECK( hipMalloc( &tree.rangeLo, cap * sizeof( int ) ) );
		HIP_CHECK( hipMalloc( &tree.rangeHi, cap * sizeof( int ) ) );
		proxyCap = cap;
		size_t need = 0;
		HIP_CHECK( hipcub::DeviceRadixSort::SortPairs( nullptr, need, codes, sortedCodes, index, sortedIndex, (int)cap, 0, 32, stream ) );
		Temp( need );
	}
};

template <typename T>
static T* GrowPinned( T*& p, size_t& cap, size_t need )
{
	if ( need > cap )
	{
		hipHostFree( p );
		cap = std::max<size_t>( 256, need + need / 2 );
		HIP_CHECK( hipHostMalloc( &p, cap * sizeof( T ), hipHostMallocDefault ) );
	}
	return p;
}

PairFinder::PairFinder()
{
	impl = new Impl();
	HIP_CHECK( hipStreamCreate( &impl->stream ) );
	HIP_CHECK( hipMalloc( &impl->bounds, 5 * sizeof( float ) ) );
	HIP_CHECK( hipMalloc( &impl->counters, 4 * sizeof( int ) ) );
	HIP_CHECK( hipHostMalloc( &impl->hostCounters, 4 * sizeof( int ), hipHostMallocDefault ) );
}

PairFinder::~PairFinder()
{
	Impl* m = impl;
	(void)hipStreamSynchronize( m->stream );
	hipFree( m->proxies );
	hipFree( m->queries );
	hipFree( m->large );
	hipFree( m->largeList );
	hipFree( m->moved );
	hipFree( m->codes );
	hipFree( m->sortedCodes );
	hipFree( m->index );
	hipFree( m->sortedIndex );
	hipFree( m->temp );
	hipFree( m->bounds );
	hipFree( m->tree.nodes );
	hipFree( m->tree.chunkBox );
	hipFree( m->tree.rightOf );
	hipFree( m->tree.rangeLo );
	hipFree( m->tree.rangeHi );
	hipFree( m->chunkBoxes );
	hipFree( m->shapeSlots );
	hipHostFree( m->hostShapeSlots );
	hipFree( m->cands );
	hipFree( m->sortedCands );
	hipFree( m->pairSlots );
	hipFree( m->pairUpdates );
	hipHostFree( m->hostPairs );
	hipFree( m->filters );
	hipHostFree( m->hostFilters );
	hipFree( m->counters );
	hipHostFree( m->hostCounters );
	hipHostFree( m->hostProxies );
	hipHostFree( m->hostQueries );
	hipHostFree( m->hostMoved );
	hipHostFree( m->hostCands );
	hipStreamDestroy( m->stream );
	delete m;
}

// Writes the (slot, key) pairs of a partial pair-set update into the mirror.
__global__ __launch_bounds__( kLbvhBlock ) void PairSetScatterKernel( const uint64_t* updates, int count, uint64_t* slots )
{
	int k = blockIdx.x * blockDim.x + threadIdx.x;
	if ( k < count )
	{
		slots[updates[2 * k]] = updates[2 * k + 1];
	}
}

void PairFinder::SyncPairSet( const uint64_t* slots, size_t capacity, const uint32_t* dirty, size_t dirtyCount, bool all )
{
	Impl* m = impl;
	hipStream_t s = m->stream;
	if ( capacity == 0 )
	{
		m->pairSize = 0;
		return;
	}
	if ( all || capacity != m->pairSize )
	{
		if ( capacity > m->pairCap )
		{
			hipFree( m->pairSlots );
			HIP_CHECK( hipMalloc( &m->pairSlots, capacity * sizeof( uint64_t ) ) );
			m->pairCap = capacity;
		}
		GrowPinned( m->hostPairs, m->hostPairCap, capacity );
		memcpy( m->hostPairs, slots, capacity * sizeof( uint64_t ) );
		HIP_CHECK( hipMemcpyAsync( m->pairSlots, m->hostPairs, capacity * sizeof( uint64_t ), hipMemcpyHostToDevice, s ) );
		m->pairSize = capacity;
		return;
	}
	if ( dirtyCount == 0 )
	{
		return;
	}
	GrowPinned( m->hostPairs, m->hostPairCap, 2 * dirtyCount );
	for ( size_t k = 0; k < dirtyCount; ++k )
	{
		m->hostPairs[2 * k] = dirty[k];
		m->hostPairs[2 * k + 1] = slots[dirty[k]];
	}
	if ( 2 * dirtyCount > m->pairUpdateCap )
	{
		hipFree( m->pairUpdates );
		m->pairUpdateCap = std::max<size_t>( 1024, 3 * dirtyCount );
		HIP_CHECK( hipMalloc( &m->pairUpdates, m->pairUpdateCap * sizeof( uint64_t ) ) );
	}
	HIP_CHECK( hipMemcpyAsync( m->pairUpdates, m->hostPairs, 2 * dirtyCount * sizeof( uint64_t ), hipMemcpyHostToDevice, s ) );
	PairSetScatterKernel<<<( (int)dirtyCount + kLbvhBlock - 1 ) / kLbvhBlock, kLbvhBlock, 0, s>>>( m->pairUpdates, (int)dirtyCount,
																								  m->pairSlots );
	HIP_CHECK( hipGetLastError() );
}

PairShapeFilter* PairFinder::FilterStaging( int count )
{
	return GrowPinned( impl->hostFilters, impl->hostFilterCap, (size_t)count );
}

void PairFinder::SyncShapeFilters( int count )
{
	Impl* m = impl;
	if ( (size_t)count > m->filterCap )
	{
		hipFree( m->filters );
		m->filterCap = std::max<size_t>( 1024, count + count / 2 );
		HIP_CHECK( hipMalloc( &m->filters, m->filterCap * sizeof( PairShapeFilter ) ) );
	}
	HIP_CHECK( hipMemcpyAsync( m->filters, m->hostFilters, count * sizeof( PairShapeFilter ), hipMemcpyHostToDevice, m->stream ) );
	m->filterCount = count;
}

LbvhProxy* PairFinder::ProxyStaging( int count )
{
	return GrowPinned( impl->hostProxies, impl->hostProxyCap, (size_t)count );
}

LbvhProxy* PairFinder::QueryStaging( int count )
{
	return GrowPinned( impl->hostQueries, impl->hostQueryCap, (size_t)count );
}

uint8_t* PairFinder::MovedStaging( int count )
{
	return GrowPinned( impl->hostMoved, impl->hostMovedCap, (size_t)count );
}

void PairFinder::SetMode( int mode )
{
	impl->mode = mode;
}

void PairFinder::Launch( int n, int nq, int movedCount )
{
	Impl* m = impl;
	hipStream_t s = m->stream;
	m->speculative = false;
	m->launchedProxies = n;
	m->launchedQueries = nq;
	if ( n == 0 || nq == 0 )
	{
		return;
	}
	m->Reserve( n );
	if ( (size_t)nq > m->queryCap )
	{
		hipFree( m->queries );
		hipFree( m->large );
		hipFree( m->largeList );
		m->queryCap = std::max<size_t>( 256, nq + nq / 2 );
		HIP_CHECK( hipMalloc( &m->queries, m->queryCap * sizeof( LbvhProxy ) ) );
		HIP_CHECK( hipMalloc( &m->large, m->queryCap ) );
		HIP_CHECK( hipMalloc( &m->largeList, m->queryCap * sizeof( int ) ) );
	}
	if ( (size_t)movedCount > m->movedCap )
	{
		hipFree( m->moved );
		m->movedCap = std::max<size_t>( 1024, movedCount + movedCount / 2 );
		HIP_CHECK( hipMalloc( &m->moved, m->movedCap ) );
	}
	if ( m->candCap == 0 )
	{
		m->candCap = std::max<size_t>( 4096, 8 * (size_t)nq );
		HIP_CHECK( hipMalloc( &m->cands, m->candCap * sizeof( uint64_t ) ) );
		HIP_CHECK( hipMalloc( &m->sortedCands, m->candCap * sizeof( uint64_t ) ) );
	}

	HIP_CHECK( hipMemcpyAsync( m->proxies, m->hostProxies, n * sizeof( LbvhProxy ), hipMemcpyHostToDevice, s ) );
	HIP_CHECK( hipMemcpyAsync( m->queries, m->hostQueries, nq * sizeof( LbvhProxy ), hipMemcpyHostToDevice, s ) );
	HIP_CHECK( hipMemcpyAsync( m->moved, m->hostMoved, movedCount, hipMemcpyHostToDevice, s ) );

	if ( m->mode == 0 )
	{
		int chunkCount = ( n + kLbvhChunk - 1 ) / kLbvhChunk;
		if ( (size_t)chunkCount > m->chunkCap )
		{
			hipFree( m->chunkBoxes );
			m->chunkCap = (size_t)chunkCount + chunkCount / 2 + 16;
			HIP_CHECK( hipMalloc( &m->chunkBoxes, m->chunkCap * sizeof( b2AABB ) ) );
		}
		int waves = kLbvhBlock / 64;
		PairChunkBoxKernel<<<( chunkCount + waves - 1 ) / waves, kLbvhBlock, 0, s>>>( m->proxies, n, m->chunkBoxes );
		HIP_CHECK( hipGetLastError() );
		m->Query();
		return;
	}

	LbvhTree t = m->tree;
	t.n = n;
	t.codes = m->sortedCodes;
	LbvhBoundsKernel<<<1, 1024, 0, s>>>( m->proxies, n, m->bounds );
	int blocks = ( n + kLbvhBlock - 1 ) / kLbvhBlock;
	LbvhMortonKernel<<<blocks, kLbvhBlock, 0, s>>>( m->proxies, n, m->bounds, m->codes, m->index );
	size_t temp = m->tempBytes;
	HIP_CHECK( hipcub::DeviceRadixSort::SortPairs( m->temp, temp, m->codes, m->sortedCodes, m->index, m->sortedIndex, n, 0, 32, s ) );
	LbvhLeafKernel<<<blocks, kLbvhBlock, 0, s>>>( m->proxies, m->sortedIndex, t );
	if ( n > 1 )
	{
		LbvhBuildKernel<<<blocks, kLbvhBlock, 0, s>>>( t );
	}
	int internalBlocks = n > 1 ? ( n - 1 + kLbvhBlock / 64 - 1 ) / ( kLbvhBlock / 64 ) : 0;
	LbvhFinishKernel<<<blocks + internalBlocks, kLbvhBlock, 0, s>>>( t, blocks );
	HIP_CHECK( hipGetLastError() );
	m->tree.n = n;
	m->Query();
}

// Query over the built tree; the candidate count and overflow flag come back asynchronously.
void PairFinder::Impl::Query()
{
	LbvhQueryArgs qa;
	qa.nodes = tree.nodes;
	qa.nodeCount = 2 * launchedProxies - 1;
	qa.queries = queries;
	qa.queryCount = launchedQueries;
	qa.moved = moved;
	qa.out = cands;
	qa.capacity = (int)candCap;
	qa.counters = counters;
	qa.bounds = bounds;
	qa.large = large;
	qa.largeList = largeList;
	qa.proxies = proxies;
	qa.proxyCount = launchedProxies;
	qa.pairSlots = pairSize > 0 ? pairSlots : nullptr;
	qa.pairMask = pairSize > 0 ? pairSize - 1 : 0;
	qa.filters = filterCount > 0 ? filters : nullptr;
	qa.queryCountOnDevice = speculative ? counters + 3 : nullptr;
	if ( speculative == false )
	{
		HIP_CHECK( hipMemsetAsync( counters, 0, 3 * sizeof( int ), stream ) );
	}
	if ( mode == 0 || speculative )
	{
		int waves = kLbvhBlock / 64;
		int chunkCount = ( launchedProxies + kLbvhChunk - 1 ) / kLbvhChunk;
		PairChunkQueryKernel<<<( launchedQueries + waves - 1 ) / waves, kLbvhBlock, 0, stream>>>( qa, chunkBoxes, chunkCount );
		HIP_CHECK( hipGetLastError() );
		HIP_CHECK( hipMemcpyAsync( hostCounters, counters, 4 * sizeof( int ), hipMemcpyDeviceToHost, stream ) );
		// the first keys travel with the count
		int eager = (int)std::min<size_t>( candCap, kEagerKeys );
		GrowPinned( hostCands, hostCandCap, (size_t)eager );
		HIP_CHECK( hipMemcpyAsync( hostCands, cands, eager * sizeof( uint64_t ), hipMemcpyDeviceToHost, stream ) );
		return;
	}
	int queryBlocks = ( launchedQueries + kLbvhBlock - 1 ) / kLbvhBlock;
	LbvhLargeListKernel<<<queryBlocks, kLbvhBlock, 0, stream>>>( qa );
	LbvhQueryKernel<<<queryBlocks, kLbvhBlock, 0, stream>>>( qa );
	LbvhLargeQueryKernel<<<( launchedProxies + kLbvhBlock - 1 ) / kLbvhBlock, kLbvhBlock, 0, stream>>>( qa );
	HIP_CHECK( hipGetLastError() );
	HIP_CHECK( hipMemcpyAsync( hostCounters, counters, 2 * sizeof( int ), hipMemcpyDeviceToHost, stream ) );
}

ShapeSlot* PairFinder::ShapeSlotStaging( int count )
{
	return GrowPinned( impl->hostShapeSlots, impl->hostShapeSlotCap, (size_t)count );
}

void PairFinder::LaunchSpeculative( int n, int shapeCount, const SolvePairInputs& in )
{
	Impl* m = impl;
	hipStream_t s = m->stream;
	m->speculative = true;
	int nq = std::max( in.capacity, 1 );
	m->launchedProxies = n;
	m->launchedQueries = nq;
	m->mode = 0;
	m->Reserve( std::max( n, 1 ) );
	if ( (size_t)nq > m->queryCap )
	{
		hipFree( m->queries );
		hipFree( m->large );
		hipFree( m->largeList );
		m->queryCap = std::max<size_t>( 256, nq + nq / 2 );
		HIP_CHECK( hipMalloc( &m->queries, m->queryCap * sizeof( LbvhProxy ) ) );
		HIP_CHECK( hipMalloc( &m->large, m->queryCap ) );
		HIP_CHECK( hipMalloc( &m->largeList, m->queryCap * sizeof( int ) ) );
	}
	int movedCount = std::max( shapeCount, 1 );
	if ( (size_t)movedCount > m->movedCap )
	{
		hipFree( m->moved );
		m->movedCap = std::max<size_t>( 1024, movedCount + movedCount / 2 );
		HIP_CHECK( hipMalloc( &m->moved, m->movedCap ) );
	}
	if ( (size_t)movedCount > m->shapeSlotCap )
	{
		hipFree( m->shapeSlots );
		m->shapeSlotCap = std::max<size_t>( 1024, movedCount + movedCount / 2 );
		HIP_CHECK( hipMalloc( &m->shapeSlots, m->shapeSlotCap * sizeof( ShapeSlot ) ) );
	}
	if ( m->candCap == 0 )
	{
		m->candCap = std::max<size_t>( 4096, 8 * (size_t)nq );
		HIP_CHECK( hipMalloc( &m->cands, m->candCap * sizeof( uint64_t ) ) );
		HIP_CHECK( hipMalloc( &m->sortedCands, m->candCap * sizeof( uint64_t ) ) );
	}
	int chunkCount = ( n + kLbvhChunk - 1 ) / kLbvhChunk;
	if ( (size_t)std::max( chunkCount, 1 ) > m->chunkCap )
	{
		hipFree( m->chunkBoxes );
		m->chunkCap = (size_t)chunkCount + chunkCount / 2 + 16;
		HIP_CHECK( hipMalloc( &m->chunkBoxes, m->chunkCap * sizeof( b2AABB ) ) );
	}

	// inputs that do not depend on the solve are copied while it runs
	if ( n > 0 )
	{
		HIP_CHECK( hipMemcpyAsync( m->proxies, m->hostProxies, n * sizeof( LbvhProxy ), hipMemcpyHostToDevice, s ) );
	}
	HIP_CHECK( hipMemcpyAsync( m->shapeSlots, m->hostShapeSlots, shapeCount * sizeof( ShapeSlot ), hipMemcpyHostToDevice, s ) );
	HIP_CHECK( hipMemsetAsync( m->moved, 0, movedCount, s ) );
	HIP_CHECK( hipMemsetAsync( m->counters, 0, 4 * sizeof( int ), s ) );

	HIP_CHECK( hipStreamWaitEvent( s, static_cast<hipEvent_t>( in.ready ), 0 ) );
	PairSpeculativeBuildKernel<<<( nq + kLbvhBlock - 1 ) / kLbvhBlock, kLbvhBlock, 0, s>>>(
		in.enlargedShapes, in.enlargedShapeAABBs, in.enlargedShapeCount, in.capacity, m->shapeSlots, shapeCount, n, m->queries, m->proxies,
		m->moved, m->counters + 3, m->counters + 1 );
	HIP_CHECK( hipGetLastError() );
	if ( n > 0 )
	{
		int waves = kLbvhBlock / 64;
		PairChunkBoxKernel<<<( chunkCount + waves - 1 ) / waves, kLbvhBlock, 0, s>>>( m->proxies, n, m->chunkBoxes );
		HIP_CHECK( hipGetLastError() );
		m->Query();
	}
	else
	{
		HIP_CHECK( hipMemcpyAsync( m->hostCounters, m->counters, 4 * sizeof( int ), hipMemcpyDeviceToHost, s ) );
	}
}

const uint64_t* PairFinder::FinishSpeculative( int* count, int* queryCount, bool* complete )
{
	Impl* m = impl;
	hipStream_t s = m->stream;
	HIP_CHECK( hipStreamSynchronize( s ) );
	m->speculative = false;
	*count = 0;
	*queryCount = m->hostCounters[3];
	int total = m->hostCounters[0];
	*complete = m->hostCounters[1] == 0 && total <= (int)m->candCap && *queryCount <= m->launchedQueries;
	if ( total > (int)m->candCap )
	{
		// the next launch gets room for this many
		hipFree( m->cands );
		hipFree( m->sortedCands );
		m->candCap = (size_t)total + total / 2;
		HIP_CHECK( hipMalloc( &m->cands, m->candCap * sizeof( uint64_t ) ) );
		HIP_CHECK( hipMalloc( &m->sortedCands, m->candCap * sizeof( uint64_t ) ) );
	}
	if ( *complete == false || total == 0 )
	{
		return nullptr;
	}
	if ( total > kEagerKeys )
	{
		GrowPinned( m->hostCands, m->hostCandCap, (size_t)total );
		HIP_CHECK( hipMemcpyAsync( m->hostCands, m->cands, total * sizeof( uint64_t ), hipMemcpyDeviceToHost, s ) );
		HIP_CHECK( hipStreamSynchronize( s ) );
	}
	*count = total;
	return m->hostCands;
}

const uint64_t* PairFinder::Finish( int* count )
{
	Impl* m = impl;
	hipStream_t s = m->stream;
	*count = 0;
	int nq = m->launchedQueries;
	if ( m->launchedProxies == 0 || nq == 0 )
	{
		return nullptr;
	}

	int total = 0;
	for ( ;; )
	{
		HIP_CHECK( hipStreamSynchronize( s ) );
		total = m->hostCounters[0];
		if ( m->hostCounters[1] != 0 )
		{
			fprintf( stderr, "box2d_amd: LBVH walk exceeded the node count (malformed tree)\n" );
			abort();
		}
		if ( total > (int)m->candCap )
		{
			hipFree( m->cands );
			hipFree( m->sortedCands );
			m->candCap = (size_t)total + total / 2;
			HIP_CHECK( hipMalloc( &m->cands, m->candCap * sizeof( uint64_t ) ) );
			HIP_CHECK( hipMalloc( &m->sortedCands, m->candCap * sizeof( uint64_t ) ) );
			m->Query();
			continue;
		}
		break;
	}
	if ( total == 0 )
	{
		return nullptr;
	}

	if ( m->mode == 0 )
	{
		// each query's keys are already in ascending proxy order
		if ( total > kEagerKeys )
		{
			GrowPinned( m->hostCands, m->hostCandCap, (size_t)total );
			HIP_CHECK( hipMemcpyAsync( m->hostCands, m->cands, total * sizeof( uint64_t ), hipMemcpyDeviceToHost, s ) );
			HIP_CHECK( hipStreamSynchronize( s ) );
		}
		*count = total;
		return m->hostCands;
	}

	// (query, proxy index) order: bits 0..31 proxy index, 32.. query index
	int endBit = 32;
	while ( endBit < 64 && ( 1ull << ( endBit - 32 ) ) < (uint64_t)nq )
	{
		++endBit;
	}
	size_t need = 0;
	HIP_CHECK( hipcub::DeviceRadixSort::SortKeys( nullptr, need, m->cands, m->sortedCands, total, 0, endBit, s ) );
	m->Temp( need );
	size_t temp2 = m->tempBytes;
	HIP_CHECK( hipcub::DeviceRadixSort::SortKeys( m->temp, temp2, m->cands, m->sortedCands, total, 0, endBit, s ) );
	GrowPinned( m->hostCands, m->hostCandCap, (size_t)total );
	HIP_CHECK( hipMemcpyAsync( m->hostCands, m->sortedCands, total * sizeof( uint64_t ), hipMemcpyDeviceToHost, s ) );
	HIP_CHECK( hipStreamSynchronize( s ) );
	*count = total;
	return m->hostCands;
}

} // namespace b2g
