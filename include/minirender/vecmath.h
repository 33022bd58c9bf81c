// vecmath.h -- the small vector / matrix / array types the minirender API
// uses (the reference takes them from the ASL library: asl::Vec2, Vec3,
// Matrix4, Array, Array2, Shared).  Same names and meaning, namespace
// minirender.  Matrix4 is row-major: m(i, j) is row i, column j; m * v is
// the affine transform ((m00 x + m01 y) + m02 z) + m03 per row, as the
// reference's Renderer.cpp:345 uses it.
//
// The containers keep the idioms the reference's callers are written in
// (samples/bench.cpp:17-48, 98; samples/main.cpp:108-112):
//   Shared<TriMesh> mesh = new TriMesh();   a raw pointer becomes shared
//   mesh->vertices << p << q;               append
//   mesh->normalsI = mesh->indices;         copies share storage (shallow)
//   a.length(), a.dup() / a.clone()         size, deep copy
#ifndef MINIRENDER_VECMATH_H
#define MINIRENDER_VECMATH_H

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstddef>
#include <initializer_list>
#include <memory>
#include <type_traits>
#include <utility>
#include <vector>

namespace minirender {

// Objects that Shared<> may wrap more than once from a raw pointer (scene
// nodes, materials): the second wrap joins the first's ownership instead of
// starting a second one.
struct RefCounted : public std::enable_shared_from_this<RefCounted> {
  virtual ~RefCounted() {}
};

// Shared ownership of a heap object, as asl::Shared: constructible from a
// raw pointer (which it then owns), from a Shared or std::shared_ptr of a
// derived type, and from nullptr.  Two Shared made from the same raw
// pointer of a RefCounted object share one count.
template <class T>
class Shared {
 public:
  Shared() = default;
  Shared(std::nullptr_t) {}
  template <class U, class = typename std::enable_if<std::is_convertible<U*, T*>::value>::type>
  Shared(U* p) : p_(adopt(p)) {}
  template <class U, class = typename std::enable_if<std::is_convertible<U*, T*>::value>::type>
  Shared(const std::shared_ptr<U>& p) : p_(p) {}
  template <class U, class = typename std::enable_if<std::is_convertible<U*, T*>::value>::type>
  Shared(const Shared<U>& p) : p_(p.shared()) {}

  T* operator->() const { return p_.get(); }
  T& operator*() const { return *p_; }
  T* get() const { return p_.get(); }
  T* ptr() const { return p_.get(); }
  explicit operator bool() const { return (bool)p_; }
  operator std::shared_ptr<T>() const { return p_; }
  const std::shared_ptr<T>& shared() const { return p_; }
  template <class U>
  bool operator==(const Shared<U>& b) const { return get() == b.get(); }
  template <class U>
  bool operator!=(const Shared<U>& b) const { return get() != b.get(); }
  bool operator==(std::nullptr_t) const { return !p_; }
  bool operator!=(std::nullptr_t) const { return (bool)p_; }

 private:
  template <class U>
  static std::shared_ptr<T> adopt(U* p) {
    if (!p) return nullptr;
    if constexpr (std::is_base_of<RefCounted, U>::value) {
      std::shared_ptr<RefCounted> owner = static_cast<RefCounted*>(p)->weak_from_this().lock();
      if (owner) return std::shared_ptr<T>(owner, static_cast<T*>(p));
    }
    return std::shared_ptr<T>(p);
  }
  std::shared_ptr<T> p_;
};

template <class T, class... A>
Shared<T> makeShared(A&&... a) {
  return Shared<T>(std::make_shared<T>(std::forward<A>(a)...));
}

// Identity of array storage: every storage gets a new id when it is created.
inline std::atomic<uint64_t> g_array_storage_ids{1};

// Dynamic array with shared storage, as asl::Array: copies and assignments
// share the elements (as the reference's per-material OBJ meshes share one
// vertex array, io.cpp:313-315); dup() and clone() make independent copies.
// The storage carries an id and a modification version: every non-const
// access (which may write) bumps the version, so a consumer that kept
// (id(), version()) knows the contents are unchanged when both still match.
// Writes through a pointer or reference obtained before that point are not
// seen.
template <class T>
class Array {
  struct Store {
    std::vector<T> v;
    uint64_t id = g_array_storage_ids.fetch_add(1, std::memory_order_relaxed);
    uint64_t version = 0;
    Store() = default;
    explicit Store(std::vector<T> x) : v(std::move(x)) {}
  };

 public:
  using value_type = T;
  using iterator = typename std::vector<T>::iterator;
  using const_iterator = typename std::vector<T>::const_iterator;

  Array() : d_(std::make_shared<Store>()) {}
  explicit Array(int n) : d_(std::make_shared<Store>(std::vector<T>((size_t)std::max(n, 0)))) {}
  Array(std::initializer_list<T> v) : d_(std::make_shared<Store>(std::vector<T>(v))) {}
  Array(const std::vector<T>& v) : d_(std::make_shared<Store>(v)) {}
  Array(std::vector<T>&& v) : d_(std::make_shared<Store>(std::move(v))) {}
  template <class It>
  Array(It first, It last) : d_(std::make_shared<Store>(std::vector<T>(first, last))) {}

  int length() const { return (int)d_->v.size(); }
  size_t size() const { return d_->v.size(); }
  bool empty() const { return d_->v.empty(); }
  const T* data() const { return d_->v.data(); }
  const T& operator[](size_t i) const { return d_->v[i]; }
  const T& back() const { return d_->v.back(); }
  const_iterator begin() const { return d_->v.begin(); }
  const_iterator end() const { return d_->v.end(); }
  // Non-const access: the caller may write.
  T* data() { return w().data(); }
  T& operator[](size_t i) { return w()[i]; }
  T& back() { return w().back(); }
  T& last() { return w().back(); }
  iterator begin() { return w().begin(); }
  iterator end() { return w().end(); }

  Array& operator<<(const T& x) { w().push_back(x); return *this; }
  void push_back(const T& x) { w().push_back(x); }
  template <class... A>
  void emplace_back(A&&... a) { w().emplace_back(std::forward<A>(a)...); }
  Array& append(const Array& b) {
    const std::vector<T> copy = b.d_->v;
    w().insert(d_->v.end(), copy.begin(), copy.end());
    return *this;
  }
  Array& resize(int n) { w().resize((size_t)std::max(n, 0)); return *this; }
  void reserve(size_t n) { d_->v.reserve(n); }
  void clear() { w().clear(); }
  void removeAt(int i) { w().erase(d_->v.begin() + i); }
  template <class It>
  void insert(iterator pos, It first, It last) { w().insert(pos, first, last); }
  void assign(size_t n, const T& v) { w().assign(n, v); }
  template <class It>
  void assign(It first, It last) { w().assign(first, last); }
  int indexOf(const T& x) const {
    for (size_t i = 0; i < d_->v.size(); i++)
      if (d_->v[i] == x) return (int)i;
    return -1;
  }

  // dup(): this array takes an independent copy of its elements (later
  // changes no longer reach arrays it shared them with); clone(): an
  // independent copy.
  Array& dup() {
    d_ = std::make_shared<Store>(d_->v);
    return *this;
  }
  Array clone() const { return Array(d_->v); }
  explicit operator bool() const { return !d_->v.empty(); }
  // Element-wise comparison.
  bool operator==(const Array& b) const { return d_ == b.d_ || d_->v == b.d_->v; }
  bool operator!=(const Array& b) const { return !(*this == b); }
  bool operator==(const std::vector<T>& b) const { return d_->v == b; }
  bool operator!=(const std::vector<T>& b) const { return d_->v != b; }

  // True when both arrays view the same storage.
  bool shares(const Array& b) const { return d_ == b.d_; }
  uint64_t id() const { return d_->id; }
  uint64_t version() const { return d_->version; }
  const std::vector<T>& vec() const { return d_->v; }
  std::vector<T>& vec() { return w(); }
  operator const std::vector<T>&() const { return d_->v; }

 private:
  std::vector<T>& w() {
    ++d_->version;
    return d_->v;
  }
  std::shared_ptr<Store> d_;
};

struct Vec2 {
  float x = 0, y = 0;
  Vec2() = default;
  Vec2(float x_, float y_) : x(x_), y(y_) {}
  Vec2 operator+(const Vec2& b) const { return Vec2(x + b.x, y + b.y); }
  Vec2 operator-(const Vec2& b) const { return Vec2(x - b.x, y - b.y); }
  Vec2 operator*(float s) const { return Vec2(x * s, y * s); }
  float operator*(const Vec2& b) const { return x * b.x + y * b.y; }  // dot
  float operator^(const Vec2& b) const { return x * b.y - y * b.x; }  // cross
  Vec2 operator/(float s) const { return Vec2(x / s, y / s); }
  Vec2 operator-() const { return Vec2(-x, -y); }
  Vec2 perpend() const { return Vec2(-y, x); }
  float length() const { return std::sqrt((*this) * (*this)); }
  Vec2 normalized() const { return *this / length(); }
};
inline Vec2 operator*(float s, const Vec2& v) { return Vec2(s * v.x, s * v.y); }

struct Vec3 {
  float x = 0, y = 0, z = 0;
  Vec3() = default;
  Vec3(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
  Vec3 operator+(const Vec3& b) const { return Vec3(x + b.x, y + b.y, z + b.z); }
  Vec3 operator-(const Vec3& b) const { return Vec3(x - b.x, y - b.y, z - b.z); }
  Vec3 operator-() const { return Vec3(-x, -y, -z); }
  Vec3 operator*(float s) const { return Vec3(x * s, y * s, z * s); }
  Vec3 operator/(float s) const { return Vec3(x / s, y / s, z / s); }
  float operator*(const Vec3& b) const { return (x * b.x + y * b.y) + z * b.z; }  // dot
  Vec3 operator^(const Vec3& b) const {  // cross
    return Vec3(y * b.z - z * b.y, z * b.x - x * b.z, x * b.y - y * b.x);
  }
  Vec3& operator+=(const Vec3& b) { x += b.x; y += b.y; z += b.z; return *this; }
  Vec3& operator-=(const Vec3& b) { x -= b.x; y -= b.y; z -= b.z; return *this; }
  Vec3& operator*=(float s) { x *= s; y *= s; z *= s; return *this; }
  Vec3& operator/=(float s) { x /= s; y /= s; z /= s; return *this; }
  float length() const { return std::sqrt((*this) * (*this)); }
  Vec3 normalized() const { return *this / length(); }
};
inline Vec3 operator*(float s, const Vec3& v) { return Vec3(s * v.x, s * v.y, s * v.z); }
inline Vec3 min(const Vec3& a, const Vec3& b) {
  return Vec3(a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z);
}
inline Vec3 max(const Vec3& a, const Vec3& b) {
  return Vec3(a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z);
}

struct Matrix4 {
  float a[4][4];
  Matrix4() { *this = identity(); }
  Matrix4(float a00, float a01, float a02, float a03, float a10, float a11, float a12,
          float a13, float a20, float a21, float a22, float a23, float a30 = 0,
          float a31 = 0, float a32 = 0, float a33 = 1) {
    const float v[16] = {a00, a01, a02, a03, a10, a11, a12, a13,
                         a20, a21, a22, a23, a30, a31, a32, a33};
    for (int i = 0; i < 16; i++) a[i / 4][i % 4] = v[i];
  }
  float& operator()(int i, int j) { return a[i][j]; }
  float operator()(int i, int j) const { return a[i][j]; }
  const float* data() const { return &a[0][0]; }

  static Matrix4 identity() {
    Matrix4 m(1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1);
    return m;
  }
  static Matrix4 translate(float x, float y, float z) {
    return Matrix4(1, 0, 0, x, 0, 1, 0, y, 0, 0, 1, z, 0, 0, 0, 1);
  }
  static Matrix4 translate(const Vec3& t) { return translate(t.x, t.y, t.z); }
  static Matrix4 scale(float s) { return Matrix4(s, 0, 0, 0, 0, s, 0, 0, 0, 0, s, 0, 0, 0, 0, 1); }
  static Matrix4 scale(const Vec3& s) {
    return Matrix4(s.x, 0, 0, 0, 0, s.y, 0, 0, 0, 0, s.z, 0, 0, 0, 0, 1);
  }
  static Matrix4 rotateX(float t) {
    const float c = std::cos(t), s = std::sin(t);
    return Matrix4(1, 0, 0, 0, 0, c, -s, 0, 0, s, c, 0, 0, 0, 0, 1);
  }
  static Matrix4 rotateY(float t) {
    const float c = std::cos(t), s = std::sin(t);
    return Matrix4(c, 0, s, 0, 0, 1, 0, 0, -s, 0, c, 0, 0, 0, 0, 1);
  }
  static Matrix4 rotateZ(float t) {
    const float c = std::cos(t), s = std::sin(t);
    return Matrix4(c, -s, 0, 0, s, c, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1);
  }
  // Rotation about `axis` by |axis| radians (a rotation vector).
  static Matrix4 rotate(const Vec3& axis) {
    const float ang = axis.length();
    if (ang == 0) return identity();
    return rotate(axis / ang, ang);
  }
  static Matrix4 rotate(const Vec3& u, float ang) {
    const float c = std::cos(ang), s = std::sin(ang), k = 1 - c;
    return Matrix4(c + u.x * u.x * k, u.x * u.y * k - u.z * s, u.x * u.z * k + u.y * s, 0,
                   u.y * u.x * k + u.z * s, c + u.y * u.y * k, u.y * u.z * k - u.x * s, 0,
                   u.z * u.x * k - u.y * s, u.z * u.y * k + u.x * s, c + u.z * u.z * k, 0,
                   0, 0, 0, 1);
  }

  Matrix4 operator*(const Matrix4& b) const {
    Matrix4 r;
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) {
        float s = 0;
        for (int k = 0; k < 4; k++) s += a[i][k] * b.a[k][j];
        r.a[i][j] = s;
      }
    return r;
  }
  // Affine transform of a point.
  Vec3 operator*(const Vec3& p) const {
    return Vec3(((a[0][0] * p.x + a[0][1] * p.y) + a[0][2] * p.z) + a[0][3],
                ((a[1][0] * p.x + a[1][1] * p.y) + a[1][2] * p.z) + a[1][3],
                ((a[2][0] * p.x + a[2][1] * p.y) + a[2][2] * p.z) + a[2][3]);
  }
  // Transform of a direction (no translation).
  Vec3 operator%(const Vec3& p) const {
    return Vec3((a[0][0] * p.x + a[0][1] * p.y) + a[0][2] * p.z,
                (a[1][0] * p.x + a[1][1] * p.y) + a[1][2] * p.z,
                (a[2][0] * p.x + a[2][1] * p.y) + a[2][2] * p.z);
  }
  Matrix4 t() const {
    Matrix4 r;
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) r.a[i][j] = a[j][i];
    return r;
  }
  Matrix4 transposed() const { return t(); }
  Matrix4 inverse() const;  // general 4x4 inverse (cofactors, in double)
};

// Row-major 2-D array (rows x cols), as asl::Array2.
template <class T>
class Array2 {
 public:
  Array2() = default;
  Array2(int rows, int cols) { resize(rows, cols); }
  void resize(int rows, int cols) {
    rows_ = rows;
    cols_ = cols;
    data_.assign((size_t)rows * (size_t)cols, T());
  }
  int rows() const { return rows_; }
  int cols() const { return cols_; }
  T& operator()(int i, int j) { return data_[(size_t)i * cols_ + j]; }
  const T& operator()(int i, int j) const { return data_[(size_t)i * cols_ + j]; }
  void set(const T& v) { std::fill(data_.begin(), data_.end(), v); }
  T* data() { return data_.data(); }
  const T* data() const { return data_.data(); }
  size_t size() const { return data_.size(); }

 private:
  int rows_ = 0, cols_ = 0;
  std::vector<T> data_;
};

}  // namespace minirender

#endif
