// db2dgrid.hpp -- drop-in host grid containers (db2dgrid.hpp:12-109 of the
// reference): same names, layout (row-major fp32, idx = y*width + x) and
// accessors, so code written against the reference compiles unchanged.  The
// index overloads accept any type with integer-convertible .x/.y members
// (glm::ivec2 in the reference) so this header needs no glm.
#pragma once

#include <algorithm>
#include <array>
#include <cassert>
#include <iomanip>
#include <iostream>
#include <utility>
#include <vector>

class Single2DGrid {
public:
  Single2DGrid(int width, int height)
      : width(width), height(height), v((size_t)width * height, 0.0f) {}
  Single2DGrid() : width(0), height(0) {}

  int idx(int x, int y) const { return y * width + x; }
  float *data() { return v.data(); }
  float const *data() const { return v.data(); }

  float &operator()(int x, int y) {
    assert(x < width && x >= 0 && y < height && y >= 0);
    return v[(size_t)idx(x, y)];
  }
  float operator()(int x, int y) const {
    assert(x < width && x >= 0 && y < height && y >= 0);
    return v[(size_t)idx(x, y)];
  }
  template <typename V, typename = decltype(std::declval<V>().x)>
  float &operator()(const V &c) { return (*this)((int)c.x, (int)c.y); }
  template <typename V, typename = decltype(std::declval<V>().x)>
  float operator()(const V &c) const { return (*this)((int)c.x, (int)c.y); }

  // mgtest/mgtest.cpp assigns scalars to whole grids (u = 0.0)
  Single2DGrid &operator=(float val) {
    fill(val);
    return *this;
  }

  void fill(float val) { std::fill(std::begin(v), std::end(v), val); }
  int width, height;

  void print() {
    for (int y = 0; y < height; y++) {
      for (int x = 0; x < width; x++)
        std::cout << std::setprecision(2) << std::setw(2) << (*this)(x, y) << " ";
      std::cout << "\n";
    }
  }

private:
  std::vector<float> v;
};

class DoubleBuffered2DGrid {
public:
  DoubleBuffered2DGrid(int width, int height)
      : width(width), height(height), front(0), back(1),
        v{std::vector<float>((size_t)width * height), std::vector<float>((size_t)width * height)} {}
  DoubleBuffered2DGrid() : width(0), height(0), front(0), back(1) {}
  DoubleBuffered2DGrid(const DoubleBuffered2DGrid &) = delete;
  DoubleBuffered2DGrid(DoubleBuffered2DGrid &&) = default;
  DoubleBuffered2DGrid &operator=(DoubleBuffered2DGrid &&) = default;

  int idx(int x, int y) const { return y * width + x; }
  void swap() { std::swap(front, back); }

  float &f(int x, int y) { return v[front][(size_t)idx(x, y)]; }
  float &b(int x, int y) { return v[back][(size_t)idx(x, y)]; }
  float f(int x, int y) const { return v[front][(size_t)idx(x, y)]; }
  float b(int x, int y) const { return v[back][(size_t)idx(x, y)]; }
  template <typename V, typename = decltype(std::declval<V>().x)>
  float &f(const V &c) { return f((int)c.x, (int)c.y); }
  template <typename V, typename = decltype(std::declval<V>().x)>
  float &b(const V &c) { return b((int)c.x, (int)c.y); }

  float &operator()(int x, int y) { return f(x, y); }
  float operator()(int x, int y) const { return f(x, y); }
  template <typename V, typename = decltype(std::declval<V>().x)>
  float &operator()(const V &c) { return f((int)c.x, (int)c.y); }

  void copyFrontToBack() { v[back] = v[front]; }

  float *data() { return v[front].data(); }
  float const *data() const { return v[front].data(); }
  float *back_data() { return v[back].data(); }
  int width, height;

private:
  int front, back;
  std::array<std::vector<float>, 2> v;
};
