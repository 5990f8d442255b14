// Walsh-Hadamard form of the Q1 element matrix.
//
// The reflections of the element box (x -> 1 - x, per axis) permute its 8 nodes and flip
// the sign of the reflected displacement component; an isotropic material's KE commutes
// with all 8 of them.  With the unnormalised 8-point Walsh-Hadamard transform H over the
// node coordinate index m = x + 2y + 4z (H_sm = (-1)^popcount(s & m), H H = 8 I), applied to
// each displacement component, the transformed entry (component i, index s) transforms
// under reflection r with the character of t = s ^ e_i (e_i: bit of axis i).  KE is
// therefore block diagonal in that basis:
//
//   KE u = H (M (H u)),   (M x)_(i, t ^ e_i) = sum_j B_t[i][j] x_(j, t ^ e_j),
//   B_t[i][j] = (H KE H)_((t ^ e_i, i), (t ^ e_j, j)) / 64,
//
// eight symmetric 3x3 blocks.  In the blocks of t = e_a the rows and columns of component a
// vanish (rigid translations), and in those of t = e_a | e_b the third component decouples:
// 33 distinct values and 45 non-zero entries, against 576 entries of KE.  An element
// product then costs two transforms (72 additions each) and the 45 block terms.
#pragma once

namespace tofe {

constexpr int kWhtVals = 33;
// For a cubic element the axis permutations make the blocks of equal shape equal: 9 values
// (diagonal and off-diagonal of t = 0 and of t = 7, of the 2 x 2 blocks of t = e_a and of
// t = e_a | e_b, and the decoupled entry of the latter).
constexpr int kWhtCubeVals = 9;
constexpr int kWhtCubeGroup[kWhtVals] = {0, 1, 1, 0, 1, 0,          // t = 0
                                         2, 3, 2, 2, 3, 2, 2, 3, 2,  // t = e_a
                                         4, 5, 4, 6, 4, 5, 4, 6, 4, 5, 4, 6,  // t = e_a | e_b
                                         7, 8, 8, 7, 8, 7};          // t = 7
// Device KE buffers hold KE (576 values), the 16 cube-orbit slots, the 33 block values and
// the 9 cube block values.
constexpr int kWhtOffset = 576 + 16;
constexpr int kWhtCubeOffset = kWhtOffset + kWhtVals;
constexpr int kKeBufLen = kWhtCubeOffset + kWhtCubeVals;

// Position of B_t[i][j] among the kWhtVals stored values, or -1 for a structural zero.
constexpr int wht_index(int t, int i, int j)
{
  if (i > j) {
    const int x = i;
    i = j;
    j = x;
  }
  const int bits = (t & 1) + ((t >> 1) & 1) + ((t >> 2) & 1);
  if (bits == 0 || bits == 3) return (bits == 0 ? 0 : 27) + i * (5 - i) / 2 + j;
  if (bits == 1) {
    const int a = t == 1 ? 0 : t == 2 ? 1 : 2;
    if (i == a || j == a) return -1;
    const int b = a == 0 ? 1 : 0;  // the two other components b < c
    return 6 + 3 * a + (i == b ? (j == b ? 0 : 1) : 2);
  }
  const int c = t == 3 ? 2 : t == 5 ? 1 : 0;  // the decoupled component
  if ((i == c) != (j == c)) return -1;
  const int base = 15 + 4 * (2 - c);
  if (i == c) return base + 3;
  const int a = c == 0 ? 1 : 0;  // the coupled components a < b
  return base + (i == a ? (j == a ? 0 : 1) : 2);
}

}  // namespace tofe
