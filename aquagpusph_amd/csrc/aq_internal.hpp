// Internal helpers shared by the HIP kernels and the C ABI launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/aquagpusph_amd.h"

namespace aq {

typedef uint32_t usize;

// ---- error reporting across the C ABI ------------------------------------
int set_error(int code, const std::string& msg);
std::string& last_error_ref();

class Error : public std::runtime_error {
  public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

  private:
    int code_;
};

#define AQ_HIP(call)                                                           \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess)                                                  \
            throw ::aq::Error(AQ_EHIP, std::string(#call " failed: ") +        \
                                           hipGetErrorString(e_));             \
    } while (0)

#define AQ_REQUIRE(cond, msg)                                                  \
    do {                                                                       \
        if (!(cond)) throw ::aq::Error(AQ_EINVAL, (msg));                      \
    } while (0)

// Link-list data of the cell-parallel walks: the sorted icell, ihoc and
// n_cells, plus the non-empty cell list (heads[k] = first particle of the
// k-th non-empty cell, n_heads[0] = their number, the walks' hand-out
// counters from n_heads[32] on,
// cell_end[c] = one past the last particle of cell c, AQ_CELL_END bits,
// with the AQ_CELL_FLUID / AQ_CELL_ELEM kind flags above them).
#define AQ_CELL_END 0x0FFFFFFFu   // cell_end bits holding the end index
#define AQ_CELL_FLUID 0x40000000u // the cell holds fluid (imove == 1)
#define AQ_CELL_ELEM 0x80000000u  // the cell holds wall elements (imove == -3)
#define AQ_CELL_OTHER 0x20000000u // the cell holds other particles (sensors,
                                  // ghosts, mirrors: imove not 1 nor -3)
// Word of the n_heads buffer holding the number of entries in the
// owner-cell list (cells flagged AQ_CELL_ELEM or AQ_CELL_OTHER, as (first
// particle, cell, end particle with the cell's flags) triplets), filled with
// the kind flags.  A cell with more than AQ_OWNER_CHUNK non-fluid particles
// (a wall edge or corner) is listed as several entries, each a run of its
// particles holding at most AQ_OWNER_CHUNK of them, so no entry takes more
// than one pass of the cooperative element walks (16 owners) and no single
// heavy cell lengthens their tail.
#define AQ_CL_NOCELLS 512u
#define AQ_OWNER_CHUNK 16u
// Capacity of the owner-cell lists (entries) for a bound `cells` on the
// non-empty cells and N particles: one entry per cell plus one per
// AQ_OWNER_CHUNK non-fluid particles past the first chunk of a cell.
inline size_t aq_owner_list_cap(size_t cells, size_t N)
{
    return cells + N / AQ_OWNER_CHUNK + 1;
}
// Word of the n_heads buffer counting the wet owner cells (owner cells with
// fluid in their neighbourhood; a second list of the same triplets).
#define AQ_CL_NWCELLS 513u
// Word of the n_heads buffer counting the near cells (cells with a fluid
// particle within the support of a wall element, listed by the fused walk).
#define AQ_CL_NNCELLS 514u
#define AQ_CL_WORDS 1024u  // size of the n_heads buffer in words
// Hand-out counters of the walks launched on the side stream beside the
// fused walk (8 counters, 32 words apart, from this word on); the walks of
// the main stream use the counters from word 32.
#define AQ_CL_SIDE 544u
struct CellArgs {
    const uint32_t* icell;
    const uint32_t* ihoc;
    const uint32_t* n_cells;
    const uint32_t* cell_end;
    const uint32_t* heads;
    const uint32_t* n_heads;
    const uint32_t* status;
    uint32_t N;
    uint32_t flagged;  // cell_end carries the kind flags
    // owner-cell list (first particle of every cell holding wall elements
    // or other non-fluid particles, counted in n_heads[AQ_CL_NOCELLS]);
    // nullptr: the element kernels visit every non-empty cell
    const uint32_t* ocells = nullptr;
    const uint32_t* wcells = nullptr;  // the wet owner cells
    const uint32_t* nlist = nullptr;   // the near cells (fused walk)
    const uint32_t* blist = nullptr;   // every non-empty cell, blocked order
    // word of the n_heads buffer where the walk's hand-out counters start
    // (0: word 32, the main stream's)
    uint32_t next_off = 0;
};

// Optional aids of the per-particle walk kernels, supplied by the runtime:
// `list` / `count` are scratch for the compacted indices of the particles
// that run a kernel (N entries and one counter), and cell_end (carrying the
// kind flags when `flagged`) lets a walk skip neighbour cells that hold no
// candidate of the wanted kind.  Kernels given none run one thread per
// particle and walk every cell.
struct PartWalk {
    uint32_t* list = nullptr;
    uint32_t* count = nullptr;
    const uint32_t* cell_end = nullptr;
    uint32_t flagged = 0;
};

// Workgroups of `block` threads of kernel `ker` resident at once on the
// current device (occupancy x compute units; cached per device).
unsigned resident_blocks(const void* ker, int block);

// Wrap a C ABI body: translate exceptions into error codes.
template <class F> inline int guard(F f, bool check_launch = true)
{
    try {
        f();
        if (!check_launch) return AQ_OK;
        hipError_t e = hipGetLastError();
        if (e != hipSuccess)
            return set_error(AQ_EHIP, std::string("kernel launch failed: ") +
                                          hipGetErrorString(e));
        return AQ_OK;
    } catch (const Error& e) {
        return set_error(e.code(), e.what());
    } catch (const std::bad_alloc&) {
        return set_error(AQ_ENOMEM, "host allocation failed");
    } catch (const std::exception& e) {
        return set_error(AQ_EINVAL, e.what());
    }
}

inline hipStream_t S(void* s) { return (hipStream_t)s; }

inline unsigned blocks_for(size_t n, unsigned bs)
{
    return (unsigned)((n + bs - 1) / bs);
}

inline int check_dims(int dims)
{
    if (dims != 2 && dims != 3)
        throw Error(AQ_EINVAL, "dims must be 2 or 3");
    return dims;
}

// Largest float s with fl(fl(sqrt(s)) / H) < 2 (the reference acceptance
// test q < SUPPORT with q = length(r_ij)/H).  sqrt and division are
// correctly rounded and monotone, so q(s) is non-decreasing in s and the
// test "q < 2" is exactly "s <= s_max" for every non-NaN s.
float support_threshold(float H);

}  // namespace aq
