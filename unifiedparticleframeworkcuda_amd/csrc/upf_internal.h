// upf_internal.h -- declarations shared by the library's translation units.
#ifndef UPF_INTERNAL_H_
#define UPF_INTERNAL_H_

#include <cstdint>
#include <vector>

#include "upf.h"

namespace upf {

uint32_t spread10(uint32_t v);
uint32_t morton3(uint32_t x, uint32_t y, uint32_t z);
bool host_key(const upf_params& p, float x, float y, float z, uint32_t& key);
float pressure_grad(float dist, float sr);
void build_luts(float sr, float* out);  // 5 * 301 floats
float sqrt_threshold(float h);
uint64_t splitmix64(uint64_t& s);
double uniform01(uint64_t& s);
int choose_domain(upf_params* p, const float box[3], float pad);
// UPF_OK when p describes a domain every device path can index: block_size a
// power of two with block_shift = log2(block_size) and grid_dim = 1024 /
// block_size (cell tables hold grid_dim^3 entries indexed by key >> 3 shift),
// cells wider than the scaled support radius, list capacity in range.
int check_params(const upf_params* p);
// Text file of `which` (UPF_TEXT_*) from host arrays; vel4, p and rho are
// read for UPF_TEXT_RESTART only.
int write_text(const char* path, int which, int n, const float* pos4, const int* type, const float* vel4,
               const float* p, const float* rho);
// Appends one record to each binary dump: n and the float4 records to
// particles_path; the number of blocks and their (block index, first, count)
// triples, sorted by block index, to blocks_path.
int write_cell_dump(const char* particles_path, const char* blocks_path, int n, const float* rec4,
                    std::vector<uint32_t>& blocks);

}  // namespace upf

#endif
