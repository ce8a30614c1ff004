// Host-side runtime pieces shared by the world bookkeeping: id pools, bit sets, assertion hooks.
//
// The id pool and bit set semantics follow the reference exactly (src/id_pool.c:19-38,
// src/bitset.h) because ids are handed out in LIFO free order and bit sets are iterated in
// ascending order, and both orders feed the deterministic contact and body ordering.
#pragma once

#include "box2d/box2d.h"
#include "box2d/constants.h"

#include <cstdint>
#include <cstring>
#include <vector>

namespace b2g
{

void Log( const char* fmt, ... );
[[noreturn]] void Fatal( const char* fmt, ... );

#define B2G_CHECK( cond )                                                                                                        \
	do                                                                                                                           \
	{                                                                                                                            \
		if ( !( cond ) )                                                                                                         \
		{                                                                                                                        \
			::b2g::Fatal( "check failed: %s (%s:%d)", #cond, __FILE__, __LINE__ );                                               \
		}                                                                                                                        \
	}                                                                                                                            \
	while ( 0 )

// LIFO id allocator: freed ids are reused most-recent first, fresh ids are handed out in order.
struct IdPool
{
	std::vector<int> freeArray;
	int nextIndex = 0;

	int Alloc()
	{
		if ( freeArray.empty() == false )
		{
			int id = freeArray.back();
			freeArray.pop_back();
			return id;
		}
		return nextIndex++;
	}

	void Free( int id )
	{
		freeArray.push_back( id );
	}

	int Count() const
	{
		return nextIndex - (int)freeArray.size();
	}

	int Capacity() const
	{
		return nextIndex;
	}
};

// Growable bit set over non-negative indices.
struct BitSet
{
	std::vector<uint64_t> words;

	void SetBitGrow( int index )
	{
		size_t w = (size_t)index >> 6;
		if ( w >= words.size() )
		{
			words.resize( w + 1 + words.size() / 2, 0 );
		}
		words[w] |= uint64_t( 1 ) << ( index & 63 );
	}

	void ClearBit( int index )
	{
		size_t w = (size_t)index >> 6;
		if ( w < words.size() )
		{
			words[w] &= ~( uint64_t( 1 ) << ( index & 63 ) );
		}
	}

	bool GetBit( int index ) const
	{
		size_t w = (size_t)index >> 6;
		if ( w >= words.size() )
		{
			return false;
		}
		return ( words[w] >> ( index & 63 ) ) & 1;
	}

	void Clear()
	{
		std::fill( words.begin(), words.end(), 0 );
	}
};

// Remove element i by moving the last element into its place. Returns the index of the element
// that moved (the old last index) or B2_NULL_INDEX if i was the last element
// (same contract as b2Array_RemoveSwap, src/container.h).
template <typename T>
inline int RemoveSwap( std::vector<T>& a, int i )
{
	int last = (int)a.size() - 1;
	if ( i == last )
	{
		a.pop_back();
		return B2_NULL_INDEX;
	}
	a[i] = a[last];
	a.pop_back();
	return last;
}

} // namespace b2g
