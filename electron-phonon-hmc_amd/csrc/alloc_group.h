// alloc_group.h -- all-or-none allocation of a group of buffers.
//
// Buffers that are only ever used together (the correlator's two propagator stores and its
// result block) are allocated as a group: when any allocation fails, the ones already made are
// released and every pointer of the group is reset to null, so a later call sees "not
// allocated" and retries instead of running on a partial set.  Plain C++ with the allocator
// passed in, so the logic is exercised on the host with a failing allocator.
#pragma once
#include <cstddef>

namespace ephmc {

// ptrs[i] receives bytes[i] bytes from alloc(void **, size_t) -> bool (true on success);
// release(void *) frees one buffer.  Returns true when every buffer was allocated.
template <typename Alloc, typename Release>
bool alloc_all_or_none(void **const *ptrs, const size_t *bytes, int n, Alloc &&alloc, Release &&release)
{
	for(int i = 0; i < n; i++) *ptrs[i] = nullptr;
	for(int i = 0; i < n; i++){
		if(!alloc(ptrs[i], bytes[i])){
			*ptrs[i] = nullptr;
			for(int j = 0; j < i; j++){
				release(*ptrs[j]);
				*ptrs[j] = nullptr;
			}
			return false;
		}
	}
	return true;
}

} // namespace ephmc
