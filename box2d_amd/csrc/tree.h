// Host-side bounding volume tree that fixes the broadphase pair order.
//
// Box2D orders new contacts by the order in which b2DynamicTree_Query visits leaves, and that
// order is a function of the tree topology: surface-area-heuristic insertion with rotations
// (reference src/dynamic_tree.c:164-673), enlarge-only refits (:798-837) and a per-step partial
// median-split rebuild (:1433-1539, :1723-1996). This class keeps that topology (and the node
// allocation order, which fixes proxy ids) identical to the reference so the pair order matches.
#pragma once

#include "core.h"

#include <cstdint>
#include <vector>

namespace b2g
{

enum TreeNodeFlags : uint16_t
{
	kNodeAllocated = 0x0001,
	kNodeEnlarged = 0x0002,
	kNodeLeaf = 0x0004,
};

struct TreeNode
{
	b2AABB aabb;
	uint64_t categoryBits;
	int child1;
	int child2;
	uint64_t userData;
	int parent; // doubles as the free-list link
	uint16_t height;
	uint16_t flags;
};

class DynamicTree
{
public:
	DynamicTree();

	int CreateProxy( b2AABB aabb, uint64_t categoryBits, uint64_t userData );
	void DestroyProxy( int proxyId );
	void MoveProxy( int proxyId, b2AABB aabb );
	void EnlargeProxy( int proxyId, b2AABB aabb );

	// EnlargeProxy with the ancestor update deferred: the leaf box is set now, and the ancestors'
	// box unions and enlarged flags are applied by FlushEnlarged, which every structural change and
	// the rebuild call first. Unions and flags do not depend on the order of enlargements, so the
	// tree ends up as with EnlargeProxy. Callers that read internal nodes (Query) flush first.
	void EnlargeProxyDeferred( int proxyId, b2AABB aabb );
	void FlushEnlarged();

	// Stack-based depth-first query; child2 is visited before child1. The callback returns false
	// to stop the query. Internal node boxes are only complete once deferred enlargements are
	// flushed, so querying with any pending is a caller error.
	template <typename F>
	void Query( b2AABB aabb, uint64_t maskBits, F&& callback ) const
	{
		if ( pendingEnlarge.empty() == false )
		{
			Fatal( "dynamic tree queried with %d deferred enlargements not flushed", (int)pendingEnlarge.size() );
		}
		if ( nodeCount == 0 )
		{
			return;
		}

		int stack[1024];
		int stackCount = 0;
		stack[stackCount++] = root;

		while ( stackCount > 0 )
		{
			int nodeId = stack[--stackCount];
			const TreeNode& node = nodes[nodeId];

			if ( b2AABB_Overlaps( node.aabb, aabb ) && ( node.categoryBits & maskBits ) != 0 )
			{
				if ( node.flags & kNodeLeaf )
				{
					if ( callback( nodeId, node.userData ) == false )
					{
						return;
					}
				}
				else if ( stackCount < 1023 )
				{
					stack[stackCount++] = node.child1;
					stack[stackCount++] = node.child2;
				}
			}
		}
	}

	// Rebuild the enlarged part of the tree (or all of it with fullBuild). Returns leaf count.
	int Rebuild( bool fullBuild );

	b2AABB GetAABB( int proxyId ) const
	{
		return nodes[proxyId].aabb;
	}

	uint64_t GetUserData( int proxyId ) const
	{
		return nodes[proxyId].userData;
	}

	int GetHeight() const
	{
		return root == B2_NULL_INDEX ? 0 : nodes[root].height;
	}

	int ProxyCount() const
	{
		return proxyCount;
	}
	// Changes whenever the tree's structure (and so its leaf order) changes; box-only changes
	// (EnlargeProxy) keep it.
	uint32_t StructureVersion() const
	{
		return structureVersion;
	}
	// Changes whenever a leaf box is set (moves and enlargements).
	uint32_t BoxVersion() const
	{
		return boxVersion;
	}

	const std::vector<TreeNode>& Nodes() const
	{
		return nodes;
	}

	// Every member in a fixed order, for world snapshots (A: the snapshot writer or reader).
	template <typename A>
	void Transfer( A& a )
	{
		a.Vec( nodes );
		a.Vec( freeStack );
		a.Pod( freshIndex );
		a.Pod( root );
		a.Pod( nodeCount );
		a.Pod( proxyCount );
		a.Pod( structureVersion );
		a.Pod( boxVersion );
		a.Vec( pendingEnlarge );
	}

	// After a restore: versions newer than any a cache may hold.
	void BumpVersions( uint32_t structureFloor, uint32_t boxFloor )
	{
		structureVersion = ( structureVersion > structureFloor ? structureVersion : structureFloor ) + 1;
		boxVersion = ( boxVersion > boxFloor ? boxVersion : boxFloor ) + 1;
	}

	int Root() const
	{
		return root;
	}

	// Calls f(proxyId, node) for every leaf in query order (the child2-first depth-first order of
	// Query). The GPU pair finder numbers proxies this way so its pairs come out in the order the
	// host query would produce them.
	template <typename F>
	void ForEachLeafInQueryOrder( F&& f ) const
	{
		if ( root == B2_NULL_INDEX )
		{
			return;
		}
		int stack[1024];
		int stackCount = 0;
		stack[stackCount++] = root;
		while ( stackCount > 0 )
		{
			int nodeId = stack[--stackCount];
			const TreeNode& node = nodes[nodeId];
			if ( node.flags & kNodeLeaf )
			{
				f( nodeId, node );
			}
			else
			{
				// at most height + 1 entries are live; the rebuild and rotations keep the height
				// logarithmic, and a deeper tree is a logic error
				if ( stackCount + 2 > 1024 )
				{
					Fatal( "dynamic tree deeper than the leaf walk supports" );
				}
				stack[stackCount++] = node.child1;
				stack[stackCount++] = node.child2;
			}
		}
	}

private:
	int AllocateNode();
	void FreeNode( int nodeId );
	int FindBestSibling( b2AABB boxD ) const;
	void RotateNodes( int iA );
	void InsertLeaf( int leaf, bool shouldRotate );
	void RemoveLeaf( int leaf );
	int BuildTree( int leafCount );

	std::vector<TreeNode> nodes;
	std::vector<int> freeStack; // freed node indices, reused last-in first-out
	int freshIndex = 0;			// next never-used node index
	int root = B2_NULL_INDEX;
	int nodeCount = 0;
	int proxyCount = 0;
	uint32_t structureVersion = 0;
	uint32_t boxVersion = 0;

	struct PendingEnlarge
	{
		int proxyId;
		b2AABB aabb;
	};
	std::vector<PendingEnlarge> pendingEnlarge;
	void EnlargeAncestors( int proxyId, b2AABB aabb );

	std::vector<int> leafIndices;
	std::vector<b2Vec2> leafCenters;
};

} // namespace b2g
