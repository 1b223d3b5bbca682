// bmi_arena_check.cpp -- C entry points over the server's host-side arena
// and admission rule (bmi_arena.h), built with g++ into
// libbmi_arena_check.so so that CPU tests can drive them without a GPU.
#include <stddef.h>

#include <vector>

#include "bmi_arena.h"

extern "C" {

void *bmi_arena_check_new(size_t capacity) { return new bmi::Arena(capacity); }
void bmi_arena_check_delete(void *a) { delete (bmi::Arena *)a; }
// returns (size_t)-1 when the request does not fit
size_t bmi_arena_check_alloc(void *a, size_t bytes, size_t align)
{
    return ((bmi::Arena *)a)->alloc(bytes, align);
}
int bmi_arena_check_release(void *a, size_t off) { return ((bmi::Arena *)a)->release(off) ? 1 : 0; }
size_t bmi_arena_check_free_bytes(void *a) { return ((bmi::Arena *)a)->free_bytes(); }
size_t bmi_arena_check_largest_free(void *a) { return ((bmi::Arena *)a)->largest_free(); }
size_t bmi_arena_check_used_blocks(void *a) { return ((bmi::Arena *)a)->used_blocks(); }
int bmi_arena_check_admit(const size_t *bytes, int n, size_t avail, int max_batch)
{
    return bmi::admit_prefix(std::vector<size_t>(bytes, bytes + n), avail, max_batch);
}
}
