// staging.h - how host threads split a chunk of a large upload.
#pragma once

#include <algorithm>
#include <cstddef>

namespace s2tc {

// Byte range [lo, hi) of a len-byte chunk that host thread t of nt copies
// into pinned staging.  Slices are whole 64-byte lines of ceil(len / nt)
// bytes rounded up, so nt of them reach at least len: together they cover
// [0, len) without overlap; trailing slices may be short or empty.
inline void stage_slice(size_t len, unsigned nt, unsigned t, size_t &lo, size_t &hi)
{
	const size_t per = ((len + nt - 1) / nt + 63) & ~(size_t)63;
	lo = std::min(len, per * t);
	hi = std::min(len, lo + per);
}

}  // namespace s2tc
