// pbrt_scene.cpp -- see pbrt_scene.h.
#include "pbrt_scene.h"

#include <array>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>

namespace pbrt_amd {

namespace {

// ------------------------------------------------------------------------------- transforms
// A 4x4 FP32 matrix with its inverse, composed as pbrt's Transform (transform.h:75-84): the
// products sum their four terms left to right, so world-space vertices come out as pbrt's.
struct Mat4 {
  float m[4][4];
};

Mat4 identity4() {
  Mat4 r{};
  for (int i = 0; i < 4; ++i) r.m[i][i] = 1.f;
  return r;
}

Mat4 mul(const Mat4& a, const Mat4& b) {
  Mat4 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      r.m[i][j] = a.m[i][0] * b.m[0][j] + a.m[i][1] * b.m[1][j] + a.m[i][2] * b.m[2][j] +
                  a.m[i][3] * b.m[3][j];
  return r;
}

Mat4 transpose(const Mat4& a) {
  Mat4 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) r.m[i][j] = a.m[j][i];
  return r;
}

// Gauss-Jordan elimination with partial pivoting, in double.  The inverse only feeds the
// camera (CameraToWorld of a Transform / ConcatTransform before Camera), never the triangles.
Mat4 inverse(const Mat4& a, bool* ok) {
  double w[4][8];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      w[i][j] = a.m[i][j];
      w[i][4 + j] = i == j ? 1.0 : 0.0;
    }
  *ok = true;
  for (int c = 0; c < 4; ++c) {
    int p = c;
    for (int r = c + 1; r < 4; ++r)
      if (std::fabs(w[r][c]) > std::fabs(w[p][c])) p = r;
    if (w[p][c] == 0.0) {
      *ok = false;
      return identity4();
    }
    if (p != c)
      for (int j = 0; j < 8; ++j) std::swap(w[p][j], w[c][j]);
    const double inv = 1.0 / w[c][c];
    for (int j = 0; j < 8; ++j) w[c][j] *= inv;
    for (int r = 0; r < 4; ++r) {
      if (r == c || w[r][c] == 0.0) continue;
      const double f = w[r][c];
      for (int j = 0; j < 8; ++j) w[r][j] -= f * w[c][j];
    }
  }
  Mat4 r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) r.m[i][j] = (float)w[i][4 + j];
  return r;
}

struct Xform {
  Mat4 m = identity4(), inv = identity4();
};

bool operator==(const Xform& a, const Xform& b) {
  return std::memcmp(&a.m, &b.m, sizeof(Mat4)) == 0 && std::memcmp(&a.inv, &b.inv, sizeof(Mat4)) == 0;
}

// a * b: b applies first (Transform::operator*).
Xform compose(const Xform& a, const Xform& b) { return {mul(a.m, b.m), mul(b.inv, a.inv)}; }

struct V3f {
  float x, y, z;
};
float length(V3f v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
V3f normalized(V3f v) {
  const float inv = 1.f / length(v);  // Vector::operator/ multiplies by the reciprocal
  return {v.x * inv, v.y * inv, v.z * inv};
}
// Cross (geometry.h:481-488): the products and differences in double, rounded to float once.
V3f cross3(V3f a, V3f b) {
  const double ax = a.x, ay = a.y, az = a.z, bx = b.x, by = b.y, bz = b.z;
  return {(float)((ay * bz) - (az * by)), (float)((az * bx) - (ax * bz)),
          (float)((ax * by) - (ay * bx))};
}

Xform translate_x(float dx, float dy, float dz) {
  Xform t;
  t.m.m[0][3] = dx;
  t.m.m[1][3] = dy;
  t.m.m[2][3] = dz;
  t.inv.m[0][3] = -dx;
  t.inv.m[1][3] = -dy;
  t.inv.m[2][3] = -dz;
  return t;
}

Xform scale_x(float x, float y, float z) {
  Xform t;
  t.m.m[0][0] = x;
  t.m.m[1][1] = y;
  t.m.m[2][2] = z;
  t.inv.m[0][0] = 1.f / x;
  t.inv.m[1][1] = 1.f / y;
  t.inv.m[2][2] = 1.f / z;
  return t;
}

// Rotation by `deg` degrees about `axis` (axis-angle matrix, transform.cpp:197-225): the angle
// goes to radians as (float)pi / 180 * deg and through sinf / cosf; the inverse is the
// transpose.
Xform rotate_x(float deg, V3f axis) {
  const V3f a = normalized(axis);
  const float rad = ((float)M_PI / 180.f) * deg;
  const float s = std::sin(rad), c = std::cos(rad);
  Mat4 r = identity4();
  r.m[0][0] = a.x * a.x + (1.f - a.x * a.x) * c;
  r.m[0][1] = a.x * a.y * (1.f - c) - a.z * s;
  r.m[0][2] = a.x * a.z * (1.f - c) + a.y * s;
  r.m[1][0] = a.x * a.y * (1.f - c) + a.z * s;
  r.m[1][1] = a.y * a.y + (1.f - a.y * a.y) * c;
  r.m[1][2] = a.y * a.z * (1.f - c) - a.x * s;
  r.m[2][0] = a.x * a.z * (1.f - c) - a.y * s;
  r.m[2][1] = a.y * a.z * (1.f - c) + a.x * s;
  r.m[2][2] = a.z * a.z + (1.f - a.z * a.z) * c;
  return {r, transpose(r)};
}

// World-from-camera frame of LookAt (transform.cpp:228-254): columns left, up, dir and the
// eye; the transform is its inverse (world to camera).
Xform lookat_x(V3f eye, V3f look, V3f up, bool* ok) {
  const V3f dir = normalized({look.x - eye.x, look.y - eye.y, look.z - eye.z});
  const V3f left = normalized(cross3(normalized(up), dir));
  const V3f nu = cross3(dir, left);
  Mat4 c2w = identity4();
  const V3f cols[3] = {left, nu, dir};
  for (int k = 0; k < 3; ++k) {
    c2w.m[0][k] = cols[k].x;
    c2w.m[1][k] = cols[k].y;
    c2w.m[2][k] = cols[k].z;
  }
  c2w.m[0][3] = eye.x;
  c2w.m[1][3] = eye.y;
  c2w.m[2][3] = eye.z;
  return {inverse(c2w, ok), c2w};
}

// Transform::operator()(const Point&) (transform.h:184-193).
Point apply_point(const Mat4& M, float x, float y, float z) {
  const auto& m = M.m;
  const float xp = m[0][0] * x + m[0][1] * y + m[0][2] * z + m[0][3];
  const float yp = m[1][0] * x + m[1][1] * y + m[1][2] * z + m[1][3];
  const float zp = m[2][0] * x + m[2][1] * y + m[2][2] * z + m[2][3];
  const float wp = m[3][0] * x + m[3][1] * y + m[3][2] * z + m[3][3];
  if (wp == 1.f) return Point(xp, yp, zp);
  const float inv = 1.f / wp;
  return Point(xp * inv, yp * inv, zp * inv);
}

// -------------------------------------------------------------------- Loop subdivision
// Loop subdivision with pbrt's rules and order (LoopSubdiv, loopsubdiv.cpp:148-485): even
// vertices by the one-ring rule (beta = 3/16 for valence 3, 3/(8 valence) otherwise, 1/16 for
// regular vertices) or the boundary rule (1/8); odd vertices 3/8, 3/8, 1/8, 1/8 (boundary
// 1/2, 1/2); every face splits into children 0..2 at its corners and child 3 in the middle;
// after the last level every vertex moves to its limit position (gamma rule, boundary 1/5).
// The new vertices of a level are the children of the old ones, in order, then the edge
// vertices in the order faces and edges first reach them; faces are the children of the old
// faces in order.  Sums are taken in pbrt's order, so the result is bit-identical.
class LoopMesh {
 public:
  LoopMesh(const std::vector<Point>& P, const std::vector<int>& vi) {
    const size_t nf = vi.size() / 3;
    V_.resize(P.size());
    for (size_t i = 0; i < P.size(); ++i) V_[i].p = {P[i].x, P[i].y, P[i].z};
    F_.resize(nf);
    for (size_t i = 0; i < nf; ++i)
      for (int j = 0; j < 3; ++j) {
        F_[i].v[j] = vi[3 * i + j];
        V_[vi[3 * i + j]].start = (int)i;  // the last face that holds the vertex
      }
    // neighbours across each edge: an edge seen a second time links the two faces
    std::map<std::pair<int, int>, std::pair<int, int>> open;
    for (size_t i = 0; i < nf; ++i)
      for (int k = 0; k < 3; ++k) {
        const int a = F_[i].v[k], b = F_[i].v[next(k)];
        const std::pair<int, int> key(std::min(a, b), std::max(a, b));
        auto it = open.find(key);
        if (it == open.end()) {
          open.emplace(key, std::make_pair((int)i, k));
        } else {
          F_[it->second.first].f[it->second.second] = (int)i;
          F_[i].f[k] = it->second.first;
          open.erase(it);
        }
      }
    for (size_t i = 0; i < V_.size(); ++i) {
      Vert& v = V_[i];
      if (v.start < 0) continue;  // a vertex no face uses
      int f = v.start;
      do {
        f = next_face(f, (int)i);
      } while (f >= 0 && f != v.start);
      v.boundary = f < 0;
      const int val = valence((int)i);
      v.regular = v.boundary ? val == 4 : val == 6;
    }
    for (size_t i = 0; i < nf; ++i) faces_.push_back((int)i);
    for (size_t i = 0; i < V_.size(); ++i) verts_.push_back((int)i);
  }

  void subdivide(int levels, std::vector<Point>* P, std::vector<int>* vi) {
    for (int l = 0; l < levels; ++l) level();
    // limit positions, computed from the last level before any is replaced
    std::vector<P3> lim(verts_.size());
    for (size_t i = 0; i < verts_.size(); ++i) {
      const int v = verts_[i];
      lim[i] = V_[v].start < 0 ? V_[v].p
               : V_[v].boundary ? weight_boundary(v, 1.f / 5.f)
                                : weight_one_ring(v, gamma(valence(v)));
    }
    std::vector<int> index(V_.size(), -1);
    P->clear();
    for (size_t i = 0; i < verts_.size(); ++i) {
      index[verts_[i]] = (int)i;
      P->push_back(Point(lim[i].x, lim[i].y, lim[i].z));
    }
    vi->clear();
    for (int f : faces_)
      for (int j = 0; j < 3; ++j) vi->push_back(index[F_[f].v[j]]);
  }

 private:
  struct P3 {
    float x, y, z;
  };
  struct Vert {
    P3 p{0.f, 0.f, 0.f};
    int start = -1, child = -1;
    bool regular = false, boundary = false;
  };
  struct Face {
    int v[3] = {-1, -1, -1}, f[3] = {-1, -1, -1}, children[4] = {-1, -1, -1, -1};
  };
  static int next(int i) { return (i + 1) % 3; }
  static int prev(int i) { return (i + 2) % 3; }
  static float beta(int valence) { return valence == 3 ? 3.f / 16.f : 3.f / (8.f * valence); }
  static float gamma(int valence) { return 1.f / (valence + 3.f / (8.f * beta(valence))); }

  int vnum(int f, int v) const {
    for (int i = 0; i < 3; ++i)
      if (F_[f].v[i] == v) return i;
    Severe("loopsubdiv: inconsistent mesh (a face does not hold its vertex)");
  }
  int next_face(int f, int v) const { return F_[f].f[vnum(f, v)]; }
  int prev_face(int f, int v) const { return F_[f].f[prev(vnum(f, v))]; }
  int next_vert(int f, int v) const { return F_[f].v[next(vnum(f, v))]; }
  int prev_vert(int f, int v) const { return F_[f].v[prev(vnum(f, v))]; }
  int other_vert(int f, int v0, int v1) const {
    for (int i = 0; i < 3; ++i)
      if (F_[f].v[i] != v0 && F_[f].v[i] != v1) return F_[f].v[i];
    Severe("loopsubdiv: inconsistent mesh (degenerate face)");
  }
  int valence(int v) const {
    int f = V_[v].start, nf = 1;
    if (!V_[v].boundary) {
      while ((f = next_face(f, v)) != V_[v].start) ++nf;
      return nf;
    }
    while ((f = next_face(f, v)) >= 0) ++nf;
    f = V_[v].start;
    while ((f = prev_face(f, v)) >= 0) ++nf;
    return nf + 1;
  }
  void one_ring(int v, std::vector<P3>* ring) const {
    ring->clear();
    int f = V_[v].start;
    if (!V_[v].boundary) {
      do {
        ring->push_back(V_[next_vert(f, v)].p);
        f = next_face(f, v);
      } while (f != V_[v].start);
      return;
    }
    for (int f2; (f2 = next_face(f, v)) >= 0;) f = f2;
    ring->push_back(V_[next_vert(f, v)].p);
    do {
      ring->push_back(V_[prev_vert(f, v)].p);
      f = prev_face(f, v);
    } while (f >= 0);
  }
  static P3 scale(float s, const P3& p) { return {s * p.x, s * p.y, s * p.z}; }
  static void add(P3* a, const P3& b) {
    a->x += b.x;
    a->y += b.y;
    a->z += b.z;
  }
  P3 weight_one_ring(int v, float b) const {
    std::vector<P3> ring;
    one_ring(v, &ring);
    const int val = (int)ring.size();
    P3 r = scale(1 - val * b, V_[v].p);
    for (const P3& q : ring) add(&r, scale(b, q));
    return r;
  }
  P3 weight_boundary(int v, float b) const {
    std::vector<P3> ring;
    one_ring(v, &ring);
    P3 r = scale(1 - 2 * b, V_[v].p);
    add(&r, scale(b, ring.front()));
    add(&r, scale(b, ring.back()));
    return r;
  }

  void level() {
    std::vector<int> nverts, nfaces;
    for (int v : verts_) {
      Vert c;
      c.regular = V_[v].regular;
      c.boundary = V_[v].boundary;
      V_[v].child = (int)V_.size();
      V_.push_back(c);
      nverts.push_back(V_[v].child);
    }
    for (int f : faces_)
      for (int k = 0; k < 4; ++k) {
        F_[f].children[k] = (int)F_.size();
        F_.push_back(Face());
        nfaces.push_back(F_[f].children[k]);
      }
    // even vertices
    for (int v : verts_) {
      if (V_[v].start < 0) continue;
      P3 q;
      if (!V_[v].boundary)
        q = weight_one_ring(v, V_[v].regular ? 1.f / 16.f : beta(valence(v)));
      else
        q = weight_boundary(v, 1.f / 8.f);
      V_[V_[v].child].p = q;
    }
    // odd vertices, one per edge
    std::map<std::pair<int, int>, int> edge_vert;
    for (int f : faces_)
      for (int k = 0; k < 3; ++k) {
        const int a = F_[f].v[k], b = F_[f].v[next(k)];
        const std::pair<int, int> key(std::min(a, b), std::max(a, b));
        if (edge_vert.count(key)) continue;
        Vert nv;
        nv.regular = true;
        nv.boundary = F_[f].f[k] < 0;
        nv.start = F_[f].children[3];
        const P3 &e0 = V_[key.first].p, &e1 = V_[key.second].p;
        if (nv.boundary) {
          nv.p = scale(0.5f, e0);
          add(&nv.p, scale(0.5f, e1));
        } else {
          nv.p = scale(3.f / 8.f, e0);
          add(&nv.p, scale(3.f / 8.f, e1));
          add(&nv.p, scale(1.f / 8.f, V_[other_vert(f, key.first, key.second)].p));
          add(&nv.p, scale(1.f / 8.f, V_[other_vert(F_[f].f[k], key.first, key.second)].p));
        }
        edge_vert[key] = (int)V_.size();
        nverts.push_back((int)V_.size());
        V_.push_back(nv);
      }
    // topology of the new level
    for (int v : verts_)
      if (V_[v].start >= 0)
        V_[V_[v].child].start = F_[V_[v].start].children[vnum(V_[v].start, v)];
    for (int f : faces_) {
      const Face& F = F_[f];
      for (int k = 0; k < 3; ++k) {
        F_[F.children[3]].f[k] = F.children[next(k)];
        F_[F.children[k]].f[next(k)] = F.children[3];
        int f2 = F.f[k];
        F_[F.children[k]].f[k] = f2 >= 0 ? F_[f2].children[vnum(f2, F.v[k])] : -1;
        f2 = F.f[prev(k)];
        F_[F.children[k]].f[prev(k)] = f2 >= 0 ? F_[f2].children[vnum(f2, F.v[k])] : -1;
      }
    }
    for (int f : faces_) {
      const Face F = F_[f];
      for (int k = 0; k < 3; ++k) {
        F_[F.children[k]].v[k] = V_[F.v[k]].child;
        const int a = F.v[k], b = F.v[next(k)];
        const int ev = edge_vert[std::make_pair(std::min(a, b), std::max(a, b))];
        F_[F.children[k]].v[next(k)] = ev;
        F_[F.children[next(k)]].v[k] = ev;
        F_[F.children[3]].v[k] = ev;
      }
    }
    verts_.swap(nverts);
    faces_.swap(nfaces);
  }

  std::vector<Vert> V_;
  std::vector<Face> F_;
  std::vector<int> verts_, faces_;  // the current level
};

// ------------------------------------------------------------------------------------ lexer
struct Token {
  enum Kind { kEnd, kNum, kStr, kIdent, kLBrack, kRBrack } kind = kEnd;
  std::string text;
  float num = 0.f;
  int line = 0;
};

// Tokens of pbrtlex.ll: numbers [-+]?(d+|d+.d*|.d+)([eE][-+]?d+)? read as (float)atof,
// quoted strings with C escapes, identifiers, brackets; '#' starts a comment to the end of
// the line.  An illegal character is reported and skipped.
class Lexer {
 public:
  Lexer(std::string text, std::string file, SceneFile* out)
      : s_(std::move(text)), file_(std::move(file)), out_(out) {}
  const std::string& file() const { return file_; }
  int line() const { return line_; }

  Token next() {
    while (pos_ < s_.size()) {
      const char c = s_[pos_];
      if (c == '\n') {
        ++line_;
        ++pos_;
      } else if (c == ' ' || c == '\t' || c == '\r') {
        ++pos_;
      } else if (c == '#') {
        while (pos_ < s_.size() && s_[pos_] != '\n') ++pos_;
      } else if (c == '[' || c == ']') {
        ++pos_;
        Token t;
        t.kind = c == '[' ? Token::kLBrack : Token::kRBrack;
        t.line = line_;
        return t;
      } else if (c == '"') {
        return string_token();
      } else if (std::isalpha((unsigned char)c) || c == '_') {
        Token t;
        t.kind = Token::kIdent;
        t.line = line_;
        while (pos_ < s_.size() && (std::isalnum((unsigned char)s_[pos_]) || s_[pos_] == '_'))
          t.text += s_[pos_++];
        return t;
      } else {
        Token t;
        if (number_token(&t)) return t;
        out_->messages.push_back(where() + "Illegal character: " + std::string(1, c));
        ++pos_;
      }
    }
    Token t;
    t.line = line_;
    return t;
  }

  std::string where() const { return file_ + ":" + std::to_string(line_) + ": "; }

 private:
  bool number_token(Token* t) {
    size_t p = pos_;
    if (p < s_.size() && (s_[p] == '+' || s_[p] == '-')) ++p;
    size_t digits = 0;
    while (p < s_.size() && std::isdigit((unsigned char)s_[p])) ++p, ++digits;
    if (p < s_.size() && s_[p] == '.') {
      ++p;
      while (p < s_.size() && std::isdigit((unsigned char)s_[p])) ++p, ++digits;
    }
    if (digits == 0) return false;
    if (p < s_.size() && (s_[p] == 'e' || s_[p] == 'E')) {
      size_t q = p + 1;
      if (q < s_.size() && (s_[q] == '+' || s_[q] == '-')) ++q;
      if (q < s_.size() && std::isdigit((unsigned char)s_[q])) {
        while (q < s_.size() && std::isdigit((unsigned char)s_[q])) ++q;
        p = q;
      }
    }
    t->kind = Token::kNum;
    t->text = s_.substr(pos_, p - pos_);
    t->num = (float)std::strtod(t->text.c_str(), nullptr);
    t->line = line_;
    pos_ = p;
    return true;
  }

  Token string_token() {
    Token t;
    t.kind = Token::kStr;
    t.line = line_;
    ++pos_;  // opening quote
    while (true) {
      if (pos_ >= s_.size()) Severe(where() + "Parsing error: unterminated string");
      const char c = s_[pos_++];
      if (c == '"') break;
      if (c == '\n') {
        out_->messages.push_back(where() + "Unterminated string!");
        ++line_;
        t.text += c;
        continue;
      }
      if (c != '\\' || pos_ >= s_.size()) {
        t.text += c;
        continue;
      }
      const char e = s_[pos_++];
      switch (e) {
        case 'n': t.text += '\n'; break;
        case 't': t.text += '\t'; break;
        case 'r': t.text += '\r'; break;
        case 'b': t.text += '\b'; break;
        case 'f': t.text += '\f'; break;
        case '\n': ++line_; break;  // line continuation
        default:
          if (std::isdigit((unsigned char)e) && pos_ + 1 < s_.size() &&
              std::isdigit((unsigned char)s_[pos_]) && std::isdigit((unsigned char)s_[pos_ + 1])) {
            int v = std::atoi(s_.substr(pos_ - 1, 3).c_str());
            while (v > 256) v -= 256;
            t.text += (char)v;
            pos_ += 2;
          } else {
            t.text += e;  // \" \\ and any other escaped character
          }
      }
    }
    return t;
  }

  std::string s_;
  std::string file_;
  SceneFile* out_;
  size_t pos_ = 0;
  int line_ = 1;
};

// ---------------------------------------------------------------------------- parameters
enum ParamType { kFloat, kInt, kBool, kPoint, kVector, kNormal, kString, kTexture, kRGB, kXYZ,
                 kBlackbody, kSpectrum };

struct Param {
  ParamType type;
  std::string name;
  std::vector<float> num;
  std::vector<std::string> str;
  std::vector<int> ints;
  std::vector<bool> bools;
};

// A typed parameter list (paramset.h): a later parameter of the same type and name replaces
// an earlier one; FindOne* only answers for single-valued parameters.
class Params {
 public:
  void add(Param p) {
    for (auto& q : list_)
      if (q.type == p.type && q.name == p.name) {
        q = std::move(p);
        return;
      }
    list_.push_back(std::move(p));
  }
  const Param* find(ParamType t, const std::string& n) const {
    for (const auto& p : list_)
      if (p.type == t && p.name == n) return &p;
    return nullptr;
  }
  float oneFloat(const std::string& n, float d) const {
    const Param* p = find(kFloat, n);
    return p && p->num.size() == 1 ? p->num[0] : d;
  }
  int oneInt(const std::string& n, int d) const {
    const Param* p = find(kInt, n);
    return p && p->ints.size() == 1 ? p->ints[0] : d;
  }
  bool oneBool(const std::string& n, bool d) const {
    const Param* p = find(kBool, n);
    return p && p->bools.size() == 1 ? (bool)p->bools[0] : d;
  }
  std::string oneString(const std::string& n, const std::string& d) const {
    const Param* p = find(kString, n);
    return p && p->str.size() == 1 ? p->str[0] : d;
  }
  const std::vector<Param>& list() const { return list_; }

 private:
  std::vector<Param> list_;
};

// Spectrum parameter ("color"/"rgb", "xyz") as its luminance y (RGBSpectrum::y, weights of
// spectrum.h); d when absent.  The renderer loop is grey.
float oneLuminance(const Params& ps, const std::string& n, float d, std::vector<std::string>* msg,
                   const std::string& where) {
  if (const Param* p = ps.find(kRGB, n))
    if (p->num.size() == 3)
      return 0.212671f * p->num[0] + 0.715160f * p->num[1] + 0.072169f * p->num[2];
  if (const Param* p = ps.find(kXYZ, n))
    if (p->num.size() == 3) return p->num[1];
  if (ps.find(kBlackbody, n) || ps.find(kSpectrum, n))
    msg->push_back(where + "spectrum \"" + n + "\" given as blackbody / sampled data: using 1");
  return d;
}

// ---------------------------------------------------------------------------------- parser
class Parser {
 public:
  Parser(SceneFile* out, std::string dir) : out_(out), dir_(std::move(dir)) {
    named_["world"][0] = named_["world"][1] = Xform();
  }

  void push(std::string text, std::string file) {
    if (lex_.size() > 32) Severe("Only 32 levels of nested Include allowed in scene files.");
    lex_.emplace_back(new Lexer(std::move(text), std::move(file), out_));
  }

  void run() {
    while (true) {
      Token t = next();
      if (t.kind == Token::kEnd) break;
      if (t.kind != Token::kIdent) syntax(t);
      directive(t);
      if (done_) break;
    }
    finish();
  }

 private:
  // ---- tokens (across Include files)
  Token next() {
    if (have_peek_) {
      have_peek_ = false;
      return peek_;
    }
    while (!lex_.empty()) {
      Token t = lex_.back()->next();
      if (t.kind != Token::kEnd) return t;
      if (lex_.size() == 1) return t;
      lex_.pop_back();
    }
    return Token();
  }
  const Token& peek() {
    if (!have_peek_) {
      peek_ = next();
      have_peek_ = true;
    }
    return peek_;
  }
  std::string where() const { return lex_.empty() ? std::string() : lex_.back()->where(); }
  [[noreturn]] void syntax(const Token& t) {
    Severe(where() + "Parsing error: syntax error, unexpected " +
           (t.kind == Token::kEnd ? std::string("end of file") : "\"" + t.text + "\""));
  }
  void msg(const std::string& m) { out_->messages.push_back(where() + m); }

  float num() {
    Token t = next();
    if (t.kind != Token::kNum) syntax(t);
    return t.num;
  }
  std::string str() {
    Token t = next();
    if (t.kind != Token::kStr) syntax(t);
    return t.text;
  }
  // num_array: "[" NUM+ "]" or one NUM
  std::vector<float> num_array() {
    std::vector<float> v;
    Token t = next();
    if (t.kind == Token::kNum) return {t.num};
    if (t.kind != Token::kLBrack) syntax(t);
    while (true) {
      t = next();
      if (t.kind == Token::kRBrack && !v.empty()) return v;
      if (t.kind != Token::kNum) syntax(t);
      v.push_back(t.num);
    }
  }

  // paramlist: ("type name" array)*, typed as InitParamSet (pbrtparse.yy:652-766)
  Params params() {
    Params ps;
    while (peek().kind == Token::kStr) {
      const std::string decl = next().text;
      std::vector<float> nums;
      std::vector<std::string> strs;
      bool is_string = false;
      Token t = next();
      if (t.kind == Token::kNum) {
        nums.push_back(t.num);
      } else if (t.kind == Token::kStr) {
        strs.push_back(t.text);
        is_string = true;
      } else if (t.kind == Token::kLBrack) {
        t = next();
        if (t.kind == Token::kStr) is_string = true;
        while (t.kind != Token::kRBrack) {
          if (is_string && t.kind == Token::kStr)
            strs.push_back(t.text);
          else if (!is_string && t.kind == Token::kNum)
            nums.push_back(t.num);
          else
            syntax(t);
          t = next();
        }
        if (nums.empty() && strs.empty()) syntax(t);
      } else {
        syntax(t);
      }
      typed(decl, std::move(nums), std::move(strs), is_string, &ps);
    }
    return ps;
  }

  void typed(const std::string& decl, std::vector<float> nums, std::vector<std::string> strs,
             bool is_string, Params* ps) {
    static const struct {
      const char* prefix;
      ParamType type;
    } kTypes[] = {{"float", kFloat},   {"integer", kInt},     {"bool", kBool},
                  {"point", kPoint},   {"vector", kVector},   {"normal", kNormal},
                  {"string", kString}, {"texture", kTexture}, {"color", kRGB},
                  {"rgb", kRGB},       {"xyz", kXYZ},         {"blackbody", kBlackbody},
                  {"spectrum", kSpectrum}};
    size_t p = 0;
    while (p < decl.size() && std::isspace((unsigned char)decl[p])) ++p;
    if (p == decl.size()) {
      msg("Parameter \"" + decl + "\" doesn't have a type declaration?!");
      return;
    }
    int k = -1;
    for (int i = 0; i < (int)(sizeof(kTypes) / sizeof(kTypes[0])); ++i)
      if (decl.compare(p, std::strlen(kTypes[i].prefix), kTypes[i].prefix) == 0) {
        k = i;
        break;
      }
    if (k < 0) {
      msg("Unable to decode type for name \"" + decl + "\"");
      return;
    }
    p += std::strlen(kTypes[k].prefix);
    while (p < decl.size() && std::isspace((unsigned char)decl[p])) ++p;
    Param q;
    q.type = kTypes[k].type;
    q.name = decl.substr(p);
    const bool wants_string = q.type == kTexture || q.type == kString || q.type == kBool;
    if (wants_string && !is_string) {
      msg("Expected string parameter value for parameter \"" + q.name + "\". Ignoring.");
      return;
    }
    if (!wants_string && q.type != kSpectrum && is_string) {
      msg("Expected numeric parameter value for parameter \"" + q.name + "\".  Ignoring.");
      return;
    }
    switch (q.type) {
      case kInt:
        for (float f : nums) q.ints.push_back((int)f);  // the grammar reads every number as float
        break;
      case kBool:
        for (const auto& s : strs) {
          if (s != "true" && s != "false")
            msg("Value \"" + s + "\" unknown for boolean parameter \"" + q.name +
                "\".Using \"false\".");
          q.bools.push_back(s == "true");
        }
        break;
      case kPoint:
      case kVector:
      case kNormal:
      case kRGB:
      case kXYZ:
        if (nums.size() % 3) {
          msg("Excess values given with parameter \"" + q.name + "\". Ignoring last " +
              std::to_string(nums.size() % 3) + " of them");
          nums.resize(nums.size() - nums.size() % 3);
        }
        q.num = std::move(nums);
        break;
      case kTexture:
        if (strs.size() != 1) {
          msg("Only one string allowed for \"texture\" parameter \"" + q.name + "\"");
          return;
        }
        q.str = std::move(strs);
        break;
      case kString:
        q.str = std::move(strs);
        break;
      default:
        q.num = std::move(nums);
        q.str = std::move(strs);
    }
    ps->add(std::move(q));
  }

  // ---- API state (api.cpp:283-305)
  bool verify_options(const char* f) {
    if (world_) msg(std::string("Options cannot be set inside world block; \"") + f +
                    "\" not allowed.  Ignoring.");
    return !world_;
  }
  bool verify_world(const char* f) {
    if (!world_) msg(std::string("Scene description must be inside world block; \"") + f +
                     "\" not allowed. Ignoring.");
    return world_;
  }
  template <class F>
  void for_active(F f) {
    for (int i = 0; i < 2; ++i)
      if (active_ & (1 << i)) cur_[i] = f(cur_[i]);
  }

  void directive(const Token& t) {
    const std::string& d = t.text;
    if (d == "Identity") {
      for_active([](const Xform&) { return Xform(); });
    } else if (d == "Translate") {
      const float x = num(), y = num(), z = num();
      for_active([&](const Xform& c) { return compose(c, translate_x(x, y, z)); });
    } else if (d == "Scale") {
      const float x = num(), y = num(), z = num();
      for_active([&](const Xform& c) { return compose(c, scale_x(x, y, z)); });
    } else if (d == "Rotate") {
      const float a = num(), x = num(), y = num(), z = num();
      for_active([&](const Xform& c) { return compose(c, rotate_x(a, {x, y, z})); });
    } else if (d == "LookAt") {
      float v[9];
      for (float& f : v) f = num();
      bool ok = true;
      const Xform l = lookat_x({v[0], v[1], v[2]}, {v[3], v[4], v[5]}, {v[6], v[7], v[8]}, &ok);
      if (!ok) msg("Singular matrix in LookAt");
      for_active([&](const Xform& c) { return compose(c, l); });
    } else if (d == "Transform" || d == "ConcatTransform") {
      const std::vector<float> a = num_array();
      if (a.size() != 16) {
        msg(d + " requires a 16 element array!");
        return;
      }
      Mat4 m;
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) m.m[i][j] = a[4 * j + i];  // the file is column-major
      bool ok = true;
      const Xform x = {m, inverse(m, &ok)};
      if (!ok) msg("Singular matrix in " + d);
      if (d == "Transform")
        for_active([&](const Xform&) { return x; });
      else
        for_active([&](const Xform& c) { return compose(c, x); });
    } else if (d == "CoordinateSystem") {
      const std::string n = str();
      named_[n][0] = cur_[0];
      named_[n][1] = cur_[1];
    } else if (d == "CoordSysTransform") {
      const std::string n = str();
      auto it = named_.find(n);
      if (it == named_.end()) {
        msg("Couldn't find named coordinate system \"" + n + "\"");
      } else {
        cur_[0] = it->second[0];
        cur_[1] = it->second[1];
      }
    } else if (d == "ActiveTransform") {
      Token w = next();
      if (w.kind != Token::kIdent) syntax(w);
      if (w.text == "All") active_ = 3;
      else if (w.text == "StartTime") active_ = 1;
      else if (w.text == "EndTime") active_ = 2;
      else syntax(w);
    } else if (d == "TransformTimes") {
      num();
      num();
      verify_options("TransformTimes");
    } else if (d == "PixelFilter" || d == "VolumeIntegrator") {
      str();
      params();
      verify_options(d.c_str());
    } else if (d == "Film") {
      str();
      Params ps = params();
      if (verify_options("Film")) film_ = std::move(ps);
    } else if (d == "Sampler") {
      std::string n = str();
      Params ps = params();
      if (verify_options("Sampler")) {
        out_->sampler = n;
        sampler_ = std::move(ps);
      }
    } else if (d == "Accelerator") {
      std::string n = str();
      Params ps = params();
      if (verify_options("Accelerator")) {
        out_->accelerator = n;
        out_->acceleratorParams = ParamSet();
        for (const Param& p : ps.list()) {
          if (p.type == kInt && p.ints.size() == 1) out_->acceleratorParams.AddInt(p.name, p.ints[0]);
          if (p.type == kBool && p.bools.size() == 1) out_->acceleratorParams.AddBool(p.name, p.bools[0]);
          if (p.type == kString && p.str.size() == 1) out_->acceleratorParams.AddString(p.name, p.str[0]);
        }
      }
    } else if (d == "SurfaceIntegrator") {
      std::string n = str();
      Params ps = params();
      if (verify_options("SurfaceIntegrator")) {
        out_->surfaceIntegrator = n;
        integrator_ = std::move(ps);
      }
    } else if (d == "Renderer") {
      std::string n = str();
      params();
      if (verify_options("Renderer")) out_->renderer = n;
    } else if (d == "Camera") {
      std::string n = str();
      Params ps = params();
      if (verify_options("Camera")) {
        out_->camera = n;
        camera_ = std::move(ps);
        // CameraToWorld = Inverse(curTransform): the stored inverse, no arithmetic
        cam_to_world_ = {cur_[0].inv, cur_[0].m};
        named_["camera"][0] = cam_to_world_;
        named_["camera"][1] = {cur_[1].inv, cur_[1].m};
      }
    } else if (d == "WorldBegin") {
      if (verify_options("WorldBegin")) {
        world_ = true;
        cur_[0] = cur_[1] = Xform();
        active_ = 3;
        named_["world"][0] = named_["world"][1] = Xform();
      }
    } else if (d == "WorldEnd") {
      if (verify_world("WorldEnd")) {
        if (!attr_stack_.empty()) msg("Missing end to pbrtAttributeBegin()");
        if (!xf_stack_.empty()) msg("Missing end to pbrtTransformBegin()");
        done_ = true;
        if (peek().kind != Token::kEnd)
          msg("directives after the first WorldEnd are not read (one scene per file)");
      }
    } else if (d == "AttributeBegin") {
      if (verify_world("AttributeBegin")) attribute_begin();
    } else if (d == "AttributeEnd") {
      if (verify_world("AttributeEnd")) attribute_end();
    } else if (d == "TransformBegin") {
      if (verify_world("TransformBegin")) xf_stack_.push_back({cur_[0], cur_[1], active_});
    } else if (d == "TransformEnd") {
      if (verify_world("TransformEnd")) {
        if (xf_stack_.empty()) {
          msg("Unmatched pbrtTransformEnd() encountered. Ignoring it.");
        } else {
          cur_[0] = xf_stack_.back().a;
          cur_[1] = xf_stack_.back().b;
          active_ = xf_stack_.back().active;
          xf_stack_.pop_back();
        }
      }
    } else if (d == "Texture") {
      str();
      str();
      str();
      params();
      verify_world("Texture");
    } else if (d == "Material" || d == "MakeNamedMaterial" || d == "AreaLightSource" ||
               d == "Volume") {
      str();
      params();
      verify_world(d.c_str());
    } else if (d == "NamedMaterial") {
      str();
      verify_world("NamedMaterial");
    } else if (d == "ReverseOrientation") {
      verify_world("ReverseOrientation");
    } else if (d == "LightSource") {
      std::string n = str();
      Params ps = params();
      if (verify_world("LightSource")) light(n, ps);
    } else if (d == "Shape") {
      std::string n = str();
      Params ps = params();
      if (verify_world("Shape")) shape(n, ps);
    } else if (d == "ObjectBegin") {
      std::string n = str();
      if (verify_world("ObjectBegin")) {
        attribute_begin();
        if (in_instance_) msg("ObjectBegin called inside of instance definition");
        in_instance_ = true;
      }
    } else if (d == "ObjectEnd") {
      if (verify_world("ObjectEnd")) {
        if (!in_instance_) msg("ObjectEnd called outside of instance definition");
        in_instance_ = false;
        attribute_end();
      }
    } else if (d == "ObjectInstance") {
      std::string n = str();
      if (verify_world("ObjectInstance")) {
        ++out_->skippedShapes;
        msg("ObjectInstance \"" + n + "\": instances are TransformedPrimitives, which the "
            "ray-space accelerators do not flatten; skipped");
      }
    } else if (d == "Include") {
      std::string n = str();
      std::string path = (!n.empty() && n[0] == '/') || dir_.empty() ? n : dir_ + "/" + n;
      std::ifstream f(path, std::ios::binary);
      if (!f) Severe("Unable to open included scene file \"" + path + "\"");
      std::stringstream ss;
      ss << f.rdbuf();
      if (have_peek_) syntax(peek_);  // cannot happen: Include's argument was just read
      push(ss.str(), path);
    } else {
      syntax(t);
    }
  }

  void attribute_begin() { attr_stack_.push_back({cur_[0], cur_[1], active_}); }
  void attribute_end() {
    if (attr_stack_.empty()) {
      msg("Unmatched pbrtAttributeEnd() encountered. Ignoring it.");
      return;
    }
    cur_[0] = attr_stack_.back().a;
    cur_[1] = attr_stack_.back().b;
    active_ = attr_stack_.back().active;
    attr_stack_.pop_back();
  }

  // LightSource "point" (point.cpp:57-64): the light sits at
  // (Translate(from) * lightToWorld)(0, 0, 0).
  void light(const std::string& n, const Params& ps) {
    if (!(cur_[0] == cur_[1]))
      msg("Animated transformations set; ignoring for \"LightSource\"and using the start "
          "transform only");
    if (n != "point") {
      msg("LightSource \"" + n + "\" is not used by the renderer loop (point lights only)");
      return;
    }
    if (out_->hasPointLight) {
      msg("more than one point light: the renderer loop uses the first");
      return;
    }
    Point from(0.f, 0.f, 0.f);
    if (const Param* p = ps.find(kPoint, "from"))
      if (p->num.size() == 3) from = Point(p->num[0], p->num[1], p->num[2]);
    const Xform l2w = compose(translate_x(from.x, from.y, from.z), cur_[0]);
    out_->render.light = apply_point(l2w.m, 0.f, 0.f, 0.f);
    const std::string w = where();
    out_->render.intensity = oneLuminance(ps, "I", 1.f, &out_->messages, w) *
                             oneLuminance(ps, "scale", 1.f, &out_->messages, w);
    out_->hasPointLight = true;
  }

  void shape(const std::string& n, const Params& ps) {
    const bool animated = !(cur_[0] == cur_[1]);
    TriangleMesh mesh;
    bool tri = false;
    if (n == "trianglemesh") {
      tri = triangle_mesh(ps, animated ? Xform() : cur_[0], &mesh);
    } else if (n == "heightfield") {
      tri = heightfield(ps, animated ? Xform() : cur_[0], &mesh);
    } else if (n == "sphere" || n == "cylinder" || n == "disk" || n == "cone" ||
               n == "paraboloid" || n == "hyperboloid") {
      ++out_->skippedShapes;  // not Triangles: the accelerators skip them
      return;
    } else if (n == "loopsubdiv") {
      tri = loop_subdiv(ps, animated ? Xform() : cur_[0], &mesh);
    } else if (n == "nurbs") {
      ++out_->skippedShapes;
      msg("Shape \"nurbs\" refines into triangles in pbrt but is not provided here; skipped");
      return;
    } else {
      msg("Shape \"" + n + "\" unknown.");
      return;
    }
    if (!tri) return;
    if (in_instance_ || animated) {
      ++out_->skippedShapes;  // TransformedPrimitive: not flattened
      if (animated) msg("Shape under an animated transform: a TransformedPrimitive; skipped");
      return;
    }
    ++nprims_;
    mesh.primitiveId = (uint32_t)nprims_;
    mesh.shapeId = nprims_;
    out_->meshes.push_back(std::move(mesh));
  }

  // CreateTriangleMeshShape (trianglemesh.cpp:363-440) + TriangleMesh's world-space vertices.
  bool triangle_mesh(const Params& ps, const Xform& o2w, TriangleMesh* m) {
    const Param* vi = ps.find(kInt, "indices");
    const Param* P = ps.find(kPoint, "P");
    const Param* uv = ps.find(kFloat, "uv");
    if (!uv) uv = ps.find(kFloat, "st");
    const size_t npi = P ? P->num.size() / 3 : 0;
    bool use_uv = uv != nullptr;
    if (uv) {
      if (uv->num.size() < 2 * npi) {
        msg("Not enough of \"uv\"s for triangle mesh.  Expencted " + std::to_string(2 * npi) +
            ", found " + std::to_string(uv->num.size()) + ".  Discarding.");
        use_uv = false;
      } else if (uv->num.size() > 2 * npi) {
        msg("More \"uv\"s provided than will be used for triangle mesh.");
      }
    }
    if (!vi || !P) return false;
    for (int v : vi->ints)
      if (v < 0 || (size_t)v >= npi) {
        msg("trianglemesh has out of-bounds vertex index " + std::to_string(v) + " (" +
            std::to_string(npi) + " \"P\" values were given");
        return false;
      }
    if (ps.find(kTexture, "alpha"))
      msg("alpha texture on a triangle mesh: not applied (every hit counts)");
    m->p.reserve(npi);
    for (size_t i = 0; i < npi; ++i)
      m->p.push_back(apply_point(o2w.m, P->num[3 * i], P->num[3 * i + 1], P->num[3 * i + 2]));
    const size_t ntri = vi->ints.size() / 3;  // TriangleMesh takes nvi / 3 triangles
    m->vi.assign(vi->ints.begin(), vi->ints.begin() + 3 * ntri);
    if (use_uv) m->uv.assign(uv->num.begin(), uv->num.begin() + 2 * npi);
    return true;
  }

  // CreateLoopSubdivShape (loopsubdiv.cpp:487-502): nlevels (3) levels of Loop subdivision of
  // the control mesh, then a triangle mesh of the limit positions (without uvs).
  bool loop_subdiv(const Params& ps, const Xform& o2w, TriangleMesh* m) {
    const Param* vi = ps.find(kInt, "indices");
    const Param* P = ps.find(kPoint, "P");
    if (!vi || !P) return false;
    const int levels = ps.oneInt("nlevels", 3);
    const size_t npi = P->num.size() / 3;
    std::vector<int> idx(vi->ints.begin(), vi->ints.begin() + 3 * (vi->ints.size() / 3));
    for (int v : idx)
      if (v < 0 || (size_t)v >= npi) {
        msg("loopsubdiv has out of-bounds vertex index " + std::to_string(v) + "; skipped");
        return false;
      }
    if (levels < 0 || levels > 8) {
      msg("loopsubdiv nlevels must be in [0, 8]; skipped");
      return false;
    }
    std::vector<Point> cp(npi), lp;
    for (size_t i = 0; i < npi; ++i) cp[i] = Point(P->num[3 * i], P->num[3 * i + 1], P->num[3 * i + 2]);
    std::vector<int> lvi;
    LoopMesh(cp, idx).subdivide(levels, &lp, &lvi);
    m->p.reserve(lp.size());
    for (const Point& q : lp) m->p.push_back(apply_point(o2w.m, q.x, q.y, q.z));
    m->vi = std::move(lvi);
    return true;
  }

  // Heightfield::Refine (heightfield.cpp:62-104): a grid over [0,1]^2 with heights Pz, two
  // triangles per cell, uv = (x, y).
  bool heightfield(const Params& ps, const Xform& o2w, TriangleMesh* m) {
    const int nu = ps.oneInt("nu", -1), nv = ps.oneInt("nv", -1);
    const Param* z = ps.find(kFloat, "Pz");
    if (nu < 2 || nv < 2 || !z || z->num.size() != (size_t)nu * nv) {
      msg("heightfield needs nu, nv >= 2 and nu * nv \"Pz\" values; skipped");
      return false;
    }
    const size_t nverts = (size_t)nu * nv;
    m->p.reserve(nverts);
    m->uv.reserve(2 * nverts);
    for (int y = 0; y < nv; ++y)
      for (int x = 0; x < nu; ++x) {
        const float px = (float)x / (float)(nu - 1), py = (float)y / (float)(nv - 1);
        m->p.push_back(apply_point(o2w.m, px, py, z->num[(size_t)y * nu + x]));
        m->uv.push_back(px);
        m->uv.push_back(py);
      }
    for (int y = 0; y < nv - 1; ++y)
      for (int x = 0; x < nu - 1; ++x) {
        const int a = x + y * nu, b = (x + 1) + y * nu, c = (x + 1) + (y + 1) * nu,
                  d = x + (y + 1) * nu;
        const int q[6] = {a, b, c, a, c, d};
        m->vi.insert(m->vi.end(), q, q + 6);
      }
    return true;
  }

  // Renderer settings from the options block, once the film size is known.
  void finish() {
    if (!world_) msg("no WorldBegin: the scene has no world block");
    RenderSettings& rs = out_->render;
    rs.xres = film_.oneInt("xresolution", 640);
    rs.yres = film_.oneInt("yresolution", 480);
    if (film_.find(kFloat, "cropwindow")) msg("Film cropwindow is not applied");
    // samples per pixel (samplers/*.cpp Create*Sampler)
    const std::string& sn = out_->sampler;
    if (sn == "stratified") {
      rs.spp = sampler_.oneInt("xsamples", 2) * sampler_.oneInt("ysamples", 2);
    } else if (sn == "adaptive") {
      rs.spp = sampler_.oneInt("maxsamples", 32);
    } else {
      rs.spp = sampler_.oneInt("pixelsamples", 4);
      if (sn == "lowdiscrepancy" && rs.spp > 0 && (rs.spp & (rs.spp - 1))) {
        msg("Pixel samples being rounded up to power of 2");
        int p = 1;
        while (p < rs.spp) p <<= 1;
        rs.spp = p;
      }
    }
    if (rs.spp < 1) rs.spp = 1;
    // perspective camera (perspective.cpp:108-145)
    if (out_->camera != "perspective")
      msg("Camera \"" + out_->camera + "\": the renderer loop provides \"perspective\" only");
    float fov = camera_.oneFloat("fov", 90.f);
    const float halffov = camera_.oneFloat("halffov", -1.f);
    if (halffov > 0.f) fov = 2.f * halffov;
    const float frame = camera_.oneFloat("frameaspectratio", float(rs.xres) / float(rs.yres));
    float sw[4];
    if (frame > 1.f) {
      sw[0] = -frame; sw[1] = frame; sw[2] = -1.f; sw[3] = 1.f;
    } else {
      sw[0] = -1.f; sw[1] = 1.f; sw[2] = -1.f / frame; sw[3] = 1.f / frame;
    }
    if (const Param* p = camera_.find(kFloat, "screenwindow")) {
      if (p->num.size() == 4)
        for (int k = 0; k < 4; ++k) sw[k] = p->num[k];
      else
        msg("\"screenwindow\" should have four values");
    }
    rs.fov_deg = fov;
    rs.has_camera_to_world = true;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 4; ++j) rs.camera_to_world[4 * i + j] = cam_to_world_.m.m[i][j];
    for (int k = 0; k < 4; ++k) rs.screen[k] = sw[k];
    rs.origin = Point(cam_to_world_.m.m[0][3], cam_to_world_.m.m[1][3], cam_to_world_.m.m[2][3]);
    // integrator: the loop shades Lambertian surfaces; path bounces up to maxdepth
    const std::string& in = out_->surfaceIntegrator;
    if (in == "path") {
      rs.max_depth = integrator_.oneInt("maxdepth", 5);
    } else {
      rs.max_depth = 0;
      if (in != "directlighting")
        msg("SurfaceIntegrator \"" + in + "\": the renderer loop runs directlighting");
    }
    if (!out_->hasPointLight) {
      msg("no point light: a unit point light at the camera is used");
      rs.light = rs.origin;
      rs.intensity = 1.f;
    }
  }

  struct Saved {
    Xform a, b;
    int active;
  };

  SceneFile* out_;
  std::string dir_;
  std::vector<std::unique_ptr<Lexer>> lex_;
  Token peek_;
  bool have_peek_ = false;
  bool world_ = false, done_ = false, in_instance_ = false;
  Xform cur_[2];
  int active_ = 3;
  std::vector<Saved> attr_stack_, xf_stack_;
  std::map<std::string, std::array<Xform, 2>> named_;
  Xform cam_to_world_;
  Params film_, sampler_, camera_, integrator_;
  int nprims_ = 0;
};

std::string dir_of(const std::string& path) {
  const size_t s = path.find_last_of('/');
  return s == std::string::npos ? std::string(".") : path.substr(0, s);
}

}  // namespace

void ParseSceneString(const std::string& text, const std::string& dir, SceneFile* out) {
  *out = SceneFile();
  Parser p(out, dir);
  p.push(text, "<string>");
  p.run();
}

void ParseSceneFile(const std::string& path, SceneFile* out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) Severe("Unable to open scene file \"" + path + "\"");
  std::stringstream ss;
  ss << f.rdbuf();
  *out = SceneFile();
  Parser p(out, dir_of(path));
  p.push(ss.str(), path);
  p.run();
}

RenderStats RenderScene(const SceneFile& scene, const std::string& accelerator_override,
                        std::vector<float>* rgb) {
  const std::string name = accelerator_override.empty() ? scene.accelerator : accelerator_override;
  const ParamSet ps = name == scene.accelerator ? scene.acceleratorParams : ParamSet();
  std::unique_ptr<Aggregate> accel(MakeAccelerator(name, scene.meshes, ps));
  return RenderBreadthFirst(*accel, scene.meshes, scene.render, rgb);
}

}  // namespace pbrt_amd
