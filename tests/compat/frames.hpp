// Length-prefixed byte frames for the drop-in header drivers: each array is
// an int64 byte count followed by its bytes.
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

struct FrameReader {
  std::vector<std::vector<char>> frames;
  size_t next = 0;
  explicit FrameReader(const char* path) {
    FILE* f = std::fopen(path, "rb");
    if (!f) {
      std::perror(path);
      std::exit(3);
    }
    int64_t n;
    while (std::fread(&n, 8, 1, f) == 1) {
      std::vector<char> b((size_t)n);
      if (n && std::fread(b.data(), 1, (size_t)n, f) != (size_t)n) std::exit(3);
      frames.push_back(std::move(b));
    }
    std::fclose(f);
  }
  // a malloc'd copy (the taco outputs and Merge's COO are realloc/free-able)
  template <class T>
  T* take(int64_t* count = nullptr, int64_t extra = 0) {
    if (next >= frames.size()) std::exit(4);
    const std::vector<char>& b = frames[next++];
    const int64_t c = (int64_t)(b.size() / sizeof(T));
    T* p = (T*)std::malloc(sizeof(T) * (size_t)(c + extra + 1));
    if (!b.empty()) std::memcpy(p, b.data(), b.size());
    if (count) *count = c;
    return p;
  }
};

struct FrameWriter {
  FILE* f;
  explicit FrameWriter(const char* path) : f(std::fopen(path, "wb")) {
    if (!f) std::exit(3);
  }
  template <class T>
  void put(const T* p, int64_t count) {
    const int64_t n = count * (int64_t)sizeof(T);
    std::fwrite(&n, 8, 1, f);
    if (n) std::fwrite(p, 1, (size_t)n, f);
  }
  ~FrameWriter() { std::fclose(f); }
};
