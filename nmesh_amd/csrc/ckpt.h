// ckpt.h -- nmesh's checkpoint variables file and CRC-64 on the host.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "mesh.h"

namespace nmx {

// reflected CRC-64, ECMA-182 polynomial, pre- and post-inverted
// (crc64_continue, crc64.c:153-165)
uint64_t crc64(uint64_t crc, const void *buf, size_t nbytes);
// CRC of the concatenation of two buffers whose CRCs started from 0, the
// second nbytes2 long (crc64_0start_combine, crc64.c:251-256)
uint64_t crc64_combine(uint64_t crc1, uint64_t crc2, uint64_t nbytes2);

// tables of the device CRC (reflected representation, bit 63 = x^0):
//   tab [8][256]: slicing-by-8, S(w) = w x^64 mod P;
//   vtab [8][256]: (byte b at position j) x^4096 mod P;
//   klane [n - 2][64], n = 2..10: x^(64 (W - 1 - last_l)) with W = 5 n^3
//     and last_l the last word t = l (mod 64) below W;
//   aw [n - 2]: (~0) x^(64 W);
//   ypow [4][256]: x^(320 d 256^j).
void crc_device_tables(std::vector<uint64_t> &tab, std::vector<uint64_t> &vtab,
                       std::vector<uint64_t> &klane, std::vector<uint64_t> &aw,
                       std::vector<uint64_t> &ypow);

// nmesh's element name "<patch>_<child octant digits>" (elmname,
// connections.c:819-866)
std::string elm_name(const Elm &E);

// variables.bin for scalarwave1's evolved variables from a register in
// element blocks [e][pi, cx, cy, cz, phi][np_e] (checkpoint_write_vl,
// checkpoint_save.c:412-468); returns "" or an error
std::string write_variables_bin(const Mesh &M, const double *blk, const std::string &path);
// read variables.bin into blk (checkpoint_load_Vars, checkpoint_load.c:
// 396-700): every element of M must be found with all five variables;
// variables of other names are skipped; returns "" or an error (blk then
// partly written)
std::string read_variables_bin(const Mesh &M, double *blk, const std::string &path);

}  // namespace nmx
