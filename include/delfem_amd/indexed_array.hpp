// delfem_amd/indexed_array.hpp -- Com::CIndexedArray, the jagged array that
// DelFEM uses for sparsity patterns (include/delfem/indexed_array.h:40-141):
// `index` holds Size()+1 row offsets, `array` the column of each entry.
// Matrix patterns hand it to AddPattern / AddMat_Dia / AddMat_NonDia.
#pragma once

#include <algorithm>
#include <vector>

namespace delfem_amd {
namespace Com {

class CIndexedArray {
 public:
  CIndexedArray() = default;
  // an empty array of `size` rows
  void InitializeSize(unsigned size) {
    index.assign(size + 1, 0);
    array.clear();
  }
  unsigned Size() const { return index.empty() ? 0u : (unsigned)index.size() - 1; }
  // the transpose of `crs` restricted to columns below `size`
  // (CIndexedArray::SetTranspose, indexed_array.h:60-83)
  void SetTranspose(unsigned size, const CIndexedArray& crs) {
    index.assign(size + 1, 0);
    for (unsigned j : crs.array)
      if (j < size) index[j + 1]++;
    for (unsigned j = 0; j < size; j++) index[j + 1] += index[j];
    array.resize(index[size]);
    std::vector<unsigned> fill(index.begin(), index.end() - 1);
    for (unsigned i = 0; i < crs.Size(); i++)
      for (unsigned q = crs.index[i]; q < crs.index[i + 1]; q++) {
        const unsigned j = crs.array[q];
        if (j < size) array[fill[j]++] = i;
      }
  }
  // every row holds every column 0..nrow-1 (CIndexedArray::Fill)
  void Fill(unsigned ncol, unsigned nrow) {
    index.resize(ncol + 1);
    array.resize((size_t)ncol * nrow);
    for (unsigned i = 0; i <= ncol; i++) index[i] = i * nrow;
    for (unsigned i = 0; i < ncol; i++)
      for (unsigned j = 0; j < nrow; j++) array[(size_t)i * nrow + j] = j;
  }
  // each row ascending (CIndexedArray::Sort)
  void Sort() {
    for (unsigned i = 0; i < Size(); i++) std::sort(array.begin() + index[i], array.begin() + index[i + 1]);
  }
  bool CheckValid() const { return !index.empty() && index.back() <= array.size(); }

  std::vector<unsigned> index;
  std::vector<unsigned> array;
};

}  // namespace Com
}  // namespace delfem_amd

#ifndef DELFEM_AMD_NO_ALIASES
namespace Com = delfem_amd::Com;
#endif
