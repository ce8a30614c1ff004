// SPDX-FileCopyrightText: 2022 Erin Catto (Box2D; insert, rotate and rebuild below restate its
// src/dynamic_tree.c so the tree topology, and with it the pair order, is the same)
// SPDX-License-Identifier: MIT
// Host-side bounding volume tree (see tree.h for why its topology must match the reference).
//
// Node allocation: the reference keeps a singly linked free list that initially threads every
// node in index order and pushes freed nodes at its head (src/dynamic_tree.c:96-147). That is
// the same as a last-in first-out stack of freed indices in front of a counter of never-used
// indices, which is what this file stores.

#include "tree.h"

#include <cfloat>

namespace b2g
{

static inline float Perimeter( b2AABB a )
{
	float wx = a.upperBound.x - a.lowerBound.x;
	float wy = a.upperBound.y - a.lowerBound.y;
	return 2.0f * ( wx + wy );
}

static inline bool EnlargeAABB( b2AABB* a, b2AABB b )
{
	bool changed = false;
	if ( b.lowerBound.x < a->lowerBound.x )
	{
		a->lowerBound.x = b.lowerBound.x;
		changed = true;
	}
	if ( b.lowerBound.y < a->lowerBound.y )
	{
		a->lowerBound.y = b.lowerBound.y;
		changed = true;
	}
	if ( a->upperBound.x < b.upperBound.x )
	{
		a->upperBound.x = b.upperBound.x;
		changed = true;
	}
	if ( a->upperBound.y < b.upperBound.y )
	{
		a->upperBound.y = b.upperBound.y;
		changed = true;
	}
	return changed;
}

static inline uint16_t MaxU16( uint16_t a, uint16_t b )
{
	return a > b ? a : b;
}

DynamicTree::DynamicTree()
{
	nodes.reserve( 64 );
}

int DynamicTree::AllocateNode()
{
	int index;
	if ( freeStack.empty() == false )
	{
		index = freeStack.back();
		freeStack.pop_back();
	}
	else
	{
		index = freshIndex++;
		if ( index >= (int)nodes.size() )
		{
			nodes.resize( index + 1 );
		}
	}

	TreeNode& node = nodes[index];
	node.aabb = { { 0.0f, 0.0f }, { 0.0f, 0.0f } };
	node.categoryBits = B2_DEFAULT_CATEGORY_BITS;
	node.child1 = B2_NULL_INDEX;
	node.child2 = B2_NULL_INDEX;
	node.userData = 0;
	node.parent = B2_NULL_INDEX;
	node.height = 0;
	node.flags = kNodeAllocated;
	nodeCount += 1;
	return index;
}

void DynamicTree::FreeNode( int nodeId )
{
	nodes[nodeId].flags = 0;
	freeStack.push_back( nodeId );
	nodeCount -= 1;
}

// Greedy descent for the cheapest sibling under the surface area heuristic
// (reference src/dynamic_tree.c:164-302).
int DynamicTree::FindBestSibling( b2AABB boxD ) const
{
	b2Vec2 centerD = b2AABB_Center( boxD );
	float areaD = Perimeter( boxD );

	b2AABB rootBox = nodes[root].aabb;
	float areaBase = Perimeter( rootBox );
	float directCost = Perimeter( b2AABB_Union( rootBox, boxD ) );
	float inheritedCost = 0.0f;

	int bestSibling = root;
	float bestCost = directCost;

	int index = root;
	while ( nodes[index].height > 0 )
	{
		int child1 = nodes[index].child1;
		int child2 = nodes[index].child2;

		float cost = directCost + inheritedCost;
		if ( cost < bestCost )
		{
			bestSibling = index;
			bestCost = cost;
		}

		inheritedCost += directCost - areaBase;

		bool leaf1 = nodes[child1].height == 0;
		bool leaf2 = nodes[child2].height == 0;

		float lowerCost1 = FLT_MAX;
		b2AABB box1 = nodes[child1].aabb;
		float directCost1 = Perimeter( b2AABB_Union( box1, boxD ) );
		float area1 = 0.0f;
		if ( leaf1 )
		{
			float cost1 = directCost1 + inheritedCost;
			if ( cost1 < bestCost )
			{
				bestSibling = child1;
				bestCost = cost1;
			}
		}
		else
		{
			area1 = Perimeter( box1 );
			lowerCost1 = inheritedCost + directCost1 + b2MinFloat( areaD - area1, 0.0f );
		}

		float lowerCost2 = FLT_MAX;
		b2AABB box2 = nodes[child2].aabb;
		float directCost2 = Perimeter( b2AABB_Union( box2, boxD ) );
		float area2 = 0.0f;
		if ( leaf2 )
		{
			float cost2 = directCost2 + inheritedCost;
			if ( cost2 < bestCost )
			{
				bestSibling = child2;
				bestCost = cost2;
			}
		}
		else
		{
			area2 = Perimeter( box2 );
			lowerCost2 = inheritedCost + directCost2 + b2MinFloat( areaD - area2, 0.0f );
		}

		if ( leaf1 && leaf2 )
		{
			break;
		}

		if ( bestCost <= lowerCost1 && bestCost <= lowerCost2 )
		{
			break;
		}

		if ( lowerCost1 == lowerCost2 && leaf1 == false )
		{
			// Tie on the area bound: fall back to center distance.
			b2Vec2 d1 = b2Sub( b2AABB_Center( box1 ), centerD );
			b2Vec2 d2 = b2Sub( b2AABB_Center( box2 ), centerD );
			lowerCost1 = b2LengthSquared( d1 );
			lowerCost2 = b2LengthSquared( d2 );
		}

		if ( lowerCost1 < lowerCost2 && leaf1 == false )
		{
			index = child1;
			areaBase = area1;
			directCost = directCost1;
		}
		else
		{
			index = child2;
			areaBase = area2;
			directCost = directCost2;
		}
	}

	return bestSibling;
}

// Tree rotation that lowers surface area (reference src/dynamic_tree.c:315-600).
// A rotation swaps one child X of node A with a grandchild Y under A's other child C; C keeps
// its remaining child Z. The bookkeeping is identical for all four swap kinds.
void DynamicTree::RotateNodes( int iA )
{
	TreeNode* A = &nodes[iA];
	if ( A->height < 2 )
	{
		return;
	}

	int iB = A->child1;
	int iC = A->child2;
	TreeNode* B = &nodes[iB];
	TreeNode* C = &nodes[iC];

	// Swap X (A's child on side sideA) with Y (grandchild of A under P on side sideP).
	auto apply = [&]( int iX, int iP, int iY, int iZ, bool xIsChild1OfA, bool yIsChild1OfP, b2AABB unionXZ ) {
		TreeNode* X = &nodes[iX];
		TreeNode* P = &nodes[iP];
		TreeNode* Y = &nodes[iY];
		TreeNode* Z = &nodes[iZ];
		if ( xIsChild1OfA )
		{
			A->child1 = iY;
		}
		else
		{
			A->child2 = iY;
		}
		if ( yIsChild1OfP )
		{
			P->child1 = iX;
		}
		else
		{
			P->child2 = iX;
		}
		X->parent = iP;
		Y->parent = iA;
		P->aabb = unionXZ;
		P->height = 1 + MaxU16( X->height, Z->height );
		A->height = 1 + MaxU16( P->height, Y->height );
		P->categoryBits = X->categoryBits | Z->categoryBits;
		A->categoryBits = P->categoryBits | Y->categoryBits;
		P->flags |= ( X->flags | Z->flags ) & kNodeEnlarged;
		A->flags |= ( P->flags | Y->flags ) & kNodeEnlarged;
	};

	if ( B->height == 0 )
	{
		// B is a leaf, C is internal with children F, G
		int iF = C->child1;
		int iG = C->child2;
		float costBase = Perimeter( C->aabb );
		b2AABB aabbBG = b2AABB_Union( B->aabb, nodes[iG].aabb );
		float costBF = Perimeter( aabbBG );
		b2AABB aabbBF = b2AABB_Union( B->aabb, nodes[iF].aabb );
		float costBG = Perimeter( aabbBF );

		if ( costBase < costBF && costBase < costBG )
		{
			return;
		}

		if ( costBF < costBG )
		{
			apply( iB, iC, iF, iG, true, true, aabbBG );
		}
		else
		{
			apply( iB, iC, iG, iF, true, false, aabbBF );
		}
	}
	else if ( C->height == 0 )
	{
		// C is a leaf, B is internal with children D, E
		int iD = B->child1;
		int iE = B->child2;
		float costBase = Perimeter( B->aabb );
		b2AABB aabbCE = b2AABB_Union( C->aabb, nodes[iE].aabb );
		float costCD = Perimeter( aabbCE );
		b2AABB aabbCD = b2AABB_Union( C->aabb, nodes[iD].aabb );
		float costCE = Perimeter( aabbCD );

		if ( costBase < costCD && costBase < costCE )
		{
			return;
		}

		if ( costCD < costCE )
		{
			apply( iC, iB, iD, iE, false, true, aabbCE );
		}
		else
		{
			apply( iC, iB, iE, iD, false, false, aabbCD );
		}
	}
	else
	{
		int iD = B->child1;
		int iE = B->child2;
		int iF = C->child1;
		int iG = C->child2;

		float areaB = Perimeter( B->aabb );
		float areaC = Perimeter( C->aabb );
		float costBase = areaB + areaC;
		int bestRotation = 0;
		float bestCost = costBase;

		b2AABB aabbBG = b2AABB_Union( B->aabb, nodes[iG].aabb );
		float costBF = areaB + Perimeter( aabbBG );
		if ( costBF < bestCost )
		{
			bestRotation = 1;
			bestCost = costBF;
		}

		b2AABB aabbBF = b2AABB_Union( B->aabb, nodes[iF].aabb );
		float costBG = areaB + Perimeter( aabbBF );
		if ( costBG < bestCost )
		{
			bestRotation = 2;
			bestCost = costBG;
		}

		b2AABB aabbCE = b2AABB_Union( C->aabb, nodes[iE].aabb );
		float costCD = areaC + Perimeter( aabbCE );
		if ( costCD < bestCost )
		{
			bestRotation = 3;
			bestCost = costCD;
		}

		b2AABB aabbCD = b2AABB_Union( C->aabb, nodes[iD].aabb );
		float costCE = areaC + Perimeter( aabbCD );
		if ( costCE < bestCost )
		{
			bestRotation = 4;
		}

		switch ( bestRotation )
		{
			case 1:
				apply( iB, iC, iF, iG, true, true, aabbBG );
				break;
			case 2:
				apply( iB, iC, iG, iF, true, false, aabbBF );
				break;
			case 3:
				apply( iC, iB, iD, iE, false, true, aabbCE );
				break;
			case 4:
				apply( iC, iB, iE, iD, false, false, aabbCD );
				break;
			default:
				break;
		}
	}
}

void DynamicTree::InsertLeaf( int leaf, bool shouldRotate )
{
	FlushEnlarged();
	structureVersion += 1;
	if ( root == B2_NULL_INDEX )
	{
		root = leaf;
		nodes[root].parent = B2_NULL_INDEX;
		return;
	}

	b2AABB leafAABB = nodes[leaf].aabb;
	int sibling = FindBestSibling( leafAABB );

	int oldParent = nodes[sibling].parent;
	int newParent = AllocateNode();

	nodes[newParent].parent = oldParent;
	nodes[newParent].aabb = b2AABB_Union( leafAABB, nodes[sibling].aabb );
	nodes[newParent].categoryBits = nodes[leaf].categoryBits | nodes[sibling].categoryBits;
	nodes[newParent].height = nodes[sibling].height + 1;
	nodes[newParent].child1 = sibling;
	nodes[newParent].child2 = leaf;
	nodes[sibling].parent = newParent;
	nodes[leaf].parent = newParent;

	if ( oldParent != B2_NULL_INDEX )
	{
		if ( nodes[oldParent].child1 == sibling )
		{
			nodes[oldParent].child1 = newParent;
		}
		else
		{
			nodes[oldParent].child2 = newParent;
		}
	}
	else
	{
		root = newParent;
	}

	int index = nodes[leaf].parent;
	while ( index != B2_NULL_INDEX )
	{
		int child1 = nodes[index].child1;
		int child2 = nodes[index].child2;

		nodes[index].aabb = b2AABB_Union( nodes[child1].aabb, nodes[child2].aabb );
		nodes[index].categoryBits = nodes[child1].categoryBits | nodes[child2].categoryBits;
		nodes[index].height = 1 + MaxU16( nodes[child1].height, nodes[child2].height );
		nodes[index].flags |= ( nodes[child1].flags | nodes[child2].flags ) & kNodeEnlarged;

		if ( shouldRotate )
		{
			RotateNodes( index );
		}

		index = nodes[index].parent;
	}
}

void DynamicTree::RemoveLeaf( int leaf )
{
	FlushEnlarged();
	structureVersion += 1;
	if ( leaf == root )
	{
		root = B2_NULL_INDEX;
		return;
	}

	int parent = nodes[leaf].parent;
	int grandParent = nodes[parent].parent;
	int sibling = nodes[parent].child1 == leaf ? nodes[parent].child2 : nodes[parent].child1;

	if ( grandParent != B2_NULL_INDEX )
	{
		if ( nodes[grandParent].child1 == parent )
		{
			nodes[grandParent].child1 = sibling;
		}
		else
		{
			nodes[grandParent].child2 = sibling;
		}
		nodes[sibling].parent = grandParent;
		FreeNode( parent );

		int index = grandParent;
		while ( index != B2_NULL_INDEX )
		{
			TreeNode& node = nodes[index];
			const TreeNode& c1 = nodes[node.child1];
			const TreeNode& c2 = nodes[node.child2];
			node.aabb = b2AABB_Union( c1.aabb, c2.aabb );
			node.categoryBits = c1.categoryBits | c2.categoryBits;
			node.height = 1 + MaxU16( c1.height, c2.height );
			index = node.parent;
		}
	}
	else
	{
		root = sibling;
		nodes[sibling].parent = B2_NULL_INDEX;
		FreeNode( parent );
	}
}

int DynamicTree::CreateProxy( b2AABB aabb, uint64_t categoryBits, uint64_t userData )
{
	int proxyId = AllocateNode();
	TreeNode& node = nodes[proxyId];
	node.aabb = aabb;
	node.userData = userData;
	node.categoryBits = categoryBits;
	node.height = 0;
	node.flags = kNodeAllocated | kNodeLeaf;

	InsertLeaf( proxyId, true );
	proxyCount += 1;
	return proxyId;
}

void DynamicTree::DestroyProxy( int proxyId )
{
	RemoveLeaf( proxyId );
	FreeNode( proxyId );
	proxyCount -= 1;
}

void DynamicTree::MoveProxy( int proxyId, b2AABB aabb )
{
	boxVersion += 1;
	RemoveLeaf( proxyId );
	nodes[proxyId].aabb = aabb;
	InsertLeaf( proxyId, false );
}

void DynamicTree::EnlargeProxy( int proxyId, b2AABB aabb )
{
	boxVersion += 1;
	FlushEnlarged();
	nodes[proxyId].aabb = aabb;
	EnlargeAncestors( proxyId, aabb );
}

void DynamicTree::EnlargeProxyDeferred( int proxyId, b2AABB aabb )
{
	boxVersion += 1;
	nodes[proxyId].aabb = aabb;
	pendingEnlarge.push_back( PendingEnlarge{ proxyId, aabb } );
}

void DynamicTree::FlushEnlarged()
{
	for ( const PendingEnlarge& p : pendingEnlarge )
	{
		EnlargeAncestors( p.proxyId, p.aabb );
	}
	pendingEnlarge.clear();
}

void DynamicTree::EnlargeAncestors( int proxyId, b2AABB aabb )
{
	int parentIndex = nodes[proxyId].parent;
	while ( parentIndex != B2_NULL_INDEX )
	{
		bool changed = EnlargeAABB( &nodes[parentIndex].aabb, aabb );
		nodes[parentIndex].flags |= kNodeEnlarged;
		parentIndex = nodes[parentIndex].parent;
		if ( changed == false )
		{
			break;
		}
	}

	while ( parentIndex != B2_NULL_INDEX )
	{
		if ( nodes[parentIndex].flags & kNodeEnlarged )
		{
			break;
		}
		nodes[parentIndex].flags |= kNodeEnlarged;
		parentIndex = nodes[parentIndex].parent;
	}
}

// Hoare partition of centers about the midpoint of the longest axis
// (reference src/dynamic_tree.c:1433-1539). Returns the split index.
static int PartitionMid( int* indices, b2Vec2* centers, int count )
{
	if ( count <= 2 )
	{
		return count / 2;
	}

	b2Vec2 lowerBound = centers[0];
	b2Vec2 upperBound = centers[0];
	for ( int i = 1; i < count; ++i )
	{
		lowerBound = b2Min( lowerBound, centers[i] );
		upperBound = b2Max( upperBound, centers[i] );
	}

	b2Vec2 d = b2Sub( upperBound, lowerBound );
	b2Vec2 c = { 0.5f * ( lowerBound.x + upperBound.x ), 0.5f * ( lowerBound.y + upperBound.y ) };

	int axis = d.x > d.y ? 0 : 1;
	float pivot = axis == 0 ? c.x : c.y;
	auto coord = [axis]( b2Vec2 v ) { return axis == 0 ? v.x : v.y; };

	int i1 = 0, i2 = count;
	while ( i1 < i2 )
	{
		while ( i1 < i2 && coord( centers[i1] ) < pivot )
		{
			i1 += 1;
		}
		while ( i1 < i2 && coord( centers[i2 - 1] ) >= pivot )
		{
			i2 -= 1;
		}
		if ( i1 < i2 )
		{
			int ti = indices[i1];
			indices[i1] = indices[i2 - 1];
			indices[i2 - 1] = ti;
			b2Vec2 tc = centers[i1];
			centers[i1] = centers[i2 - 1];
			centers[i2 - 1] = tc;
			i1 += 1;
			i2 -= 1;
		}
	}

	if ( i1 > 0 && i1 < count )
	{
		return i1;
	}
	return count / 2;
}

// Top-down median build over leafIndices[0, leafCount) with an explicit stack
// (reference src/dynamic_tree.c:1723-1876).
int DynamicTree::BuildTree( int leafCount )
{
	if ( leafCount == 1 )
	{
		nodes[leafIndices[0]].parent = B2_NULL_INDEX;
		return leafIndices[0];
	}

	struct Item
	{
		int nodeIndex;
		int childCount;
		int startIndex;
		int splitIndex;
		int endIndex;
	};

	std::vector<Item> stack( 1024 );
	int top = 0;

	stack[0].nodeIndex = AllocateNode();
	stack[0].childCount = -1;
	stack[0].startIndex = 0;
	stack[0].endIndex = leafCount;
	stack[0].splitIndex = PartitionMid( leafIndices.data(), leafCenters.data(), leafCount );

	while ( true )
	{
		Item* item = &stack[top];
		item->childCount += 1;

		if ( item->childCount == 2 )
		{
			if ( top == 0 )
			{
				break;
			}

			Item* parentItem = &stack[top - 1];
			TreeNode& parentNode = nodes[parentItem->nodeIndex];
			if ( parentItem->childCount == 0 )
			{
				parentNode.child1 = item->nodeIndex;
			}
			else
			{
				parentNode.child2 = item->nodeIndex;
			}

			TreeNode& node = nodes[item->nodeIndex];
			node.parent = parentItem->nodeIndex;
			const TreeNode& c1 = nodes[node.child1];
			const TreeNode& c2 = nodes[node.child2];
			node.aabb = b2AABB_Union( c1.aabb, c2.aabb );
			node.height = 1 + MaxU16( c1.height, c2.height );
			node.categoryBits = c1.categoryBits | c2.categoryBits;
			top -= 1;
		}
		else
		{
			int startIndex, endIndex;
			if ( item->childCount == 0 )
			{
				startIndex = item->startIndex;
				endIndex = item->splitIndex;
			}
			else
			{
				startIndex = item->splitIndex;
				endIndex = item->endIndex;
			}

			int count = endIndex - startIndex;
			if ( count == 1 )
			{
				int childIndex = leafIndices[startIndex];
				TreeNode& node = nodes[item->nodeIndex];
				if ( item->childCount == 0 )
				{
					node.child1 = childIndex;
				}
				else
				{
					node.child2 = childIndex;
				}
				nodes[childIndex].parent = item->nodeIndex;
			}
			else
			{
				top += 1;
				B2G_CHECK( top < 1024 );
				// AllocateNode may grow the node array, so refresh pointers afterwards.
				int newNode = AllocateNode();
				Item* newItem = &stack[top];
				newItem->nodeIndex = newNode;
				newItem->childCount = -1;
				newItem->startIndex = startIndex;
				newItem->endIndex = endIndex;
				newItem->splitIndex =
					PartitionMid( leafIndices.data() + startIndex, leafCenters.data() + startIndex, count ) + startIndex;
			}
		}
	}

	TreeNode& rootNode = nodes[stack[0].nodeIndex];
	const TreeNode& c1 = nodes[rootNode.child1];
	const TreeNode& c2 = nodes[rootNode.child2];
	rootNode.aabb = b2AABB_Union( c1.aabb, c2.aabb );
	rootNode.height = 1 + MaxU16( c1.height, c2.height );
	rootNode.categoryBits = c1.categoryBits | c2.categoryBits;
	return stack[0].nodeIndex;
}

int DynamicTree::Rebuild( bool fullBuild )
{
	FlushEnlarged();
	structureVersion += 1;
	if ( proxyCount == 0 )
	{
		return 0;
	}

	if ( (int)leafIndices.size() < proxyCount )
	{
		leafIndices.resize( proxyCount + proxyCount / 2 );
		leafCenters.resize( proxyCount + proxyCount / 2 );
	}

	int leafCount = 0;
	int stack[1024];
	int stackCount = 0;

	int nodeIndex = root;
	while ( true )
	{
		TreeNode& node = nodes[nodeIndex];
		if ( node.height == 0 || ( ( node.flags & kNodeEnlarged ) == 0 && fullBuild == false ) )
		{
			leafIndices[leafCount] = nodeIndex;
			leafCenters[leafCount] = b2AABB_Center( node.aabb );
			leafCount += 1;
			node.parent = B2_NULL_INDEX;
		}
		else
		{
			int doomed = nodeIndex;
			nodeIndex = node.child1;
			if ( stackCount < 1024 )
			{
				stack[stackCount++] = node.child2;
			}
			FreeNode( doomed );
			continue;
		}

		if ( stackCount == 0 )
		{
			break;
		}
		nodeIndex = stack[--stackCount];
	}

	root = BuildTree( leafCount );
	return leafCount;
}

} // namespace b2g
