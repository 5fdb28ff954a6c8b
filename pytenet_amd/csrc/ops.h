// Internal interfaces between the kernel files of the sweep library.
#pragma once
#include <memory>
#include "common.h"

namespace ptn {

struct Dims { int64_t Dl, d, Dr, Ml, Mr; };

// `batch` independent problems of identical shape, every tensor stacked
// contiguously along a leading batch axis (batch 1 = the single problem)
size_t heff_ws_bytes(int prec, const Dims& D, int64_t batch = 1);
size_t env_ws_bytes(int prec, const Dims& D, int64_t batch = 1);
size_t bond_ws_bytes(int prec, int64_t Dl, int64_t Dr, int64_t M, int64_t batch = 1);

int heff_apply(int prec, const Dims& D, const void* l, int tl, const void* w, const void* r, int tr,
               const void* a, int ta, void* out, int tout, void* ws, size_t ws_bytes,
               hipStream_t s, int64_t batch = 1);
int bond_apply(int prec, int64_t Dl, int64_t Dr, int64_t M, const void* l, int tl, const void* r,
               int tr, const void* c, int tcin, void* out, int tout, void* ws, size_t ws_bytes,
               hipStream_t s, int64_t batch = 1);
int env_right(int prec, const Dims& D, const void* a, const void* b, const void* w, const void* r,
              int tr, void* out, void* ws, size_t ws_bytes, hipStream_t s, int64_t batch = 1);
int env_left(int prec, const Dims& D, const void* a, const void* b, const void* w, const void* l,
             int tl, void* out, void* ws, size_t ws_bytes, hipStream_t s, int64_t batch = 1);
int convert_c128_c64(int64_t n, const void* x, void* y, hipStream_t s);

// sector-grouped H_eff (U(1) quantum numbers, batch 1): the plan of an
// operator, or null when its grouped products do not apply (no or unsuitable
// quantum numbers); built and uploaded synchronously on first use, so look it
// up outside stream capture
struct SectorPlan;
using SectorPlanRef = std::shared_ptr<const SectorPlan>;   // keeps a plan alive while in use
enum { SECTORS_HEFF = 0, SECTORS_ENV_RIGHT = 1, SECTORS_ENV_LEFT = 2 };
int sector_plan_get(const ptn_local_op* op, int kind, SectorPlanRef* out);
int heff_apply_sectors(int prec, const Dims& D, const SectorPlan* sp, const void* l, int tl,
                       const void* w, const void* r, int tr, const void* a, void* out, void* ws,
                       size_t ws_bytes, hipStream_t s);
// environment steps (site operator op: Dl, d, Dr, Ml, Mr, W and labels);
// workspace of env_ws_bytes
int env_right_sectors(int prec, const Dims& D, const SectorPlan* sp, const void* a, const void* b,
                      const void* w, const void* r, int tr, void* out, void* ws, size_t ws_bytes,
                      hipStream_t s);
int env_left_sectors(int prec, const Dims& D, const SectorPlan* sp, const void* a, const void* b,
                     const void* w, const void* l, int tl, void* out, void* ws, size_t ws_bytes,
                     hipStream_t s);

}  // namespace ptn
