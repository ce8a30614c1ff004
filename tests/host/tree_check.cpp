// Host check of the broadphase tree: enlarging proxies with the ancestor update deferred
// (EnlargeProxyDeferred, flushed by the next structural change or rebuild) leaves the same tree
// as enlarging them at once (EnlargeProxy), in any enlargement order, through partial and full
// rebuilds and proxy churn, and queries made after the flush visit the same leaves in the same
// order. Prints one FNV-1a hash of every query hit and every node per variant.
#include "tree.h"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace b2g
{
void Fatal( const char* fmt, ... )
{
	va_list a;
	va_start( a, fmt );
	vfprintf( stderr, fmt, a );
	va_end( a );
	abort();
}
} // namespace b2g

using namespace b2g;

static uint32_t s_rng = 12345u;
static float Random()
{
	s_rng ^= s_rng << 13;
	s_rng ^= s_rng >> 17;
	s_rng ^= s_rng << 5;
	return (float)( s_rng % 100000u ) / 100000.0f;
}

static uint64_t Run( int n, int variant )
{
	s_rng = 12345u;
	DynamicTree tree;
	std::vector<int> ids;
	uint64_t h = 1469598103934665603ull;
	auto mix = [&]( const void* p, size_t s ) {
		const unsigned char* c = (const unsigned char*)p;
		for ( size_t i = 0; i < s; ++i )
		{
			h ^= c[i];
			h *= 1099511628211ull;
		}
	};
	for ( int i = 0; i < n; ++i )
	{
		float x = 200.0f * Random(), y = 100.0f * Random(), w = 0.2f + Random(), h = 0.2f + Random();
		ids.push_back( tree.CreateProxy( b2AABB{ { x, y }, { x + w, y + h } }, 1, (uint64_t)i ) );
	}
	for ( int round = 0; round < 8; ++round )
	{
		std::vector<std::pair<int, b2AABB>> moves;
		for ( int k = 0; k < n; k += 1 + round % 3 )
		{
			b2AABB a = tree.GetAABB( ids[k] );
			float dx = 3.0f * ( Random() - 0.5f ), dy = 3.0f * ( Random() - 0.5f );
			moves.push_back( { ids[k], b2AABB{ { a.lowerBound.x + dx - 0.1f, a.lowerBound.y + dy - 0.1f },
											   { a.upperBound.x + dx + 0.1f, a.upperBound.y + dy + 0.1f } } } );
		}
		if ( variant == 2 )
		{
			std::reverse( moves.begin(), moves.end() );
		}
		for ( const auto& m : moves )
		{
			if ( variant == 0 )
			{
				tree.EnlargeProxy( m.first, m.second );
			}
			else
			{
				tree.EnlargeProxyDeferred( m.first, m.second );
			}
		}
		if ( variant != 0 )
		{
			tree.FlushEnlarged();
		}
		// queries between the refit and the rebuild (the pair query's position in a step) visit
		// the same leaves in the same order after a flush as after immediate enlargement
		for ( int k = 0; k < n; k += 53 )
		{
			b2AABB box = tree.GetAABB( ids[k] );
			tree.Query( box, ~0ull, [&]( int proxyId, uint64_t userData ) {
				mix( &proxyId, sizeof( proxyId ) );
				mix( &userData, sizeof( userData ) );
				return true;
			} );
		}
		if ( round == 4 )
		{
			for ( int k = 0; k < n; k += 37 )
			{
				tree.DestroyProxy( ids[k] );
				ids[k] = tree.CreateProxy( b2AABB{ { 1.0f, 1.0f }, { 2.0f, 2.0f } }, 1, (uint64_t)k );
			}
		}
		tree.Rebuild( round == 7 );
	}
	for ( const TreeNode& node : tree.Nodes() )
	{
		mix( &node.aabb, sizeof( node.aabb ) );
		mix( &node.child1, sizeof( node.child1 ) );
		mix( &node.child2, sizeof( node.child2 ) );
		mix( &node.parent, sizeof( node.parent ) );
		mix( &node.height, sizeof( node.height ) );
		mix( &node.flags, sizeof( node.flags ) );
	}
	int root = tree.Root();
	mix( &root, sizeof( root ) );
	return h;
}

int main( int argc, char** argv )
{
	int n = argc > 1 ? atoi( argv[1] ) : 2000;
	printf( "%016llx %016llx %016llx\n", (unsigned long long)Run( n, 0 ), (unsigned long long)Run( n, 1 ),
			(unsigned long long)Run( n, 2 ) );
	return 0;
}
