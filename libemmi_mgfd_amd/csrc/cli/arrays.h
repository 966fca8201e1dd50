// arrays.h -- the pointer-table arrays the carriers hold: a[i3][i2][i1] over
// one contiguous block whose base is &a[0][0][0] (the layout of
// src/cstd.c:263-330), zero-filled.
#pragma once

#include <cstdlib>
#include <cstring>

#include "emmi_cli.h"

namespace emmi_cli {

inline void* zalloc(size_t bytes) {
  void* p = calloc(bytes ? bytes : 1, 1);
  if (!p) emmi_cli_err("out of host memory (%zu bytes)", bytes);
  return p;
}

template <class T>
T* alloc1(size_t n1) {
  return static_cast<T*>(zalloc(n1 * sizeof(T)));
}

template <class T>
T** alloc2(size_t n1, size_t n2) {
  T** p = static_cast<T**>(zalloc(n2 * sizeof(T*)));
  p[0] = alloc1<T>(n1 * n2);
  for (size_t i = 1; i < n2; ++i) p[i] = p[0] + i * n1;
  return p;
}

template <class T>
T*** alloc3(size_t n1, size_t n2, size_t n3) {
  T*** p = static_cast<T***>(zalloc(n3 * sizeof(T**)));
  p[0] = static_cast<T**>(zalloc(n3 * n2 * sizeof(T*)));
  p[0][0] = alloc1<T>(n1 * n2 * n3);
  for (size_t k = 0; k < n3; ++k) {
    p[k] = p[0] + k * n2;
    for (size_t j = 0; j < n2; ++j) p[k][j] = p[0][0] + (k * n2 + j) * n1;
  }
  return p;
}

template <class T>
T**** alloc4(size_t n1, size_t n2, size_t n3, size_t n4) {
  T**** p = static_cast<T****>(zalloc(n4 * sizeof(T***)));
  p[0] = static_cast<T***>(zalloc(n4 * n3 * sizeof(T**)));
  p[0][0] = static_cast<T**>(zalloc(n4 * n3 * n2 * sizeof(T*)));
  p[0][0][0] = alloc1<T>(n1 * n2 * n3 * n4);
  for (size_t l = 0; l < n4; ++l) {
    p[l] = p[0] + l * n3;
    for (size_t k = 0; k < n3; ++k) {
      p[l][k] = p[0][0] + (l * n3 + k) * n2;
      for (size_t j = 0; j < n2; ++j) p[l][k][j] = p[0][0][0] + ((l * n3 + k) * n2 + j) * n1;
    }
  }
  return p;
}

template <class T>
void free2(T** p) {
  if (!p) return;
  free(p[0]);
  free(p);
}

template <class T>
void free3(T*** p) {
  if (!p) return;
  free(p[0][0]);
  free(p[0]);
  free(p);
}

template <class T>
void free4(T**** p) {
  if (!p) return;
  free(p[0][0][0]);
  free(p[0][0]);
  free(p[0]);
  free(p);
}

}  // namespace emmi_cli
