// Scene files: the reference's .dumb scene description (Dson text) and the
// Wavefront OBJ meshes it names, turned into the host arrays the rest of the
// library takes (dr_octree_build triangles, dr_scene_set_normals and
// dr_scene_set_materials inputs, dr_material / dr_light tables, and the camera
// and settings of dr_render_params, plus textures and texture coordinates).
//
// Behaviour follows DumbRenderContext.cu (fromFile :310-335, fromDson :337-371,
// groups / items / sub-groups :106-271, camera :476-514, renderer :515-640,
// include :641-659, materials / lights / objects :662-787), Dson.cu:224-447
// (parser), MeshReader.impl.h (OBJ reader), PolyMesh.impl.h:140-152 (fan
// triangulation), BakedTriMesh.impl.h:16-18 and Transform.impl.h (transforms),
// and the fromDson of each registered plugin (DumbRenderContextRegistry.cu).
// Arithmetic is fp32 evaluated as written (the library is compiled without
// contraction); sin, cos, tan and sqrt are the C library's single-precision ones.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/dumbray.h"

namespace dr {
int fail(int code, const std::string &msg);
bool png_decode_file(const std::string &path, std::vector<unsigned char> &rgba, uint32_t &w, uint32_t &h,
                     std::string &err);
}

namespace {

// ------------------------------------------------------------------ fp32 vectors
// Vector3.impl.h: dot = x*x' + y*y' + z*z' left to right, cross as operator&,
// v / f = v * (1 / f), normalized = v * (1 / sqrt(x*x + y*y + z*z)).
struct F3 {
  float x, y, z;
};
inline F3 f3(float x, float y, float z) { return F3{x, y, z}; }
inline F3 add(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline F3 sub(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline F3 scale(F3 a, float f) { return f3(a.x * f, a.y * f, a.z * f); }
inline F3 hadamard(F3 a, F3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }   // operator^
inline F3 divide(F3 a, F3 b) { return f3(a.x / b.x, a.y / b.y, a.z / b.z); }     // operator/(Vector3)
inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 cross(F3 u, F3 v) { return f3(u.y * v.z - v.y * u.z, v.x * u.z - u.x * v.z, u.x * v.y - v.x * u.y); }
inline F3 normalized(F3 v) {
  const float inv = 1.0f / std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
  return scale(v, inv);
}

constexpr float kRadian = 57.29577951308233f;   // Vector3.h:9

// ------------------------------------------------------------------ Transform
// Transform.impl.h: rows trans.x/y/z start as the unit axes and are turned by the
// Euler angles (degrees; y, then x, then z, each a rotation of the other two rows
// about the third, :258-262), then scaled component-wise (:274-276).  A point maps
// to (v . trans.x + p.x, ...) (:85-87); a direction to
// d.x * (trans.x / scale) + d.y * (trans.y / scale) + d.z * (trans.z / scale) (:166-168).
struct Transform {
  F3 pos{0.0f, 0.0f, 0.0f}, euler{0.0f, 0.0f, 0.0f}, scl{1.0f, 1.0f, 1.0f};
  F3 tx{}, ty{}, tz{};
  Transform() { rebake(); }

  // Vector3::rotateTwoAxisAgainstThisRadian (Vector3.impl.h:458-459, :511-513)
  static void turn_pair(F3 axis, F3 &v1, F3 &v2, float angle) {
    const float s = std::sin(angle), c = std::cos(angle);
    F3 k = cross(axis, v1);
    v1 = f3(v1.x * c - k.x * s, v1.y * c - k.y * s, v1.z * c - k.z * s);
    k = cross(axis, v2);
    v2 = f3(v2.x * c - k.x * s, v2.y * c - k.y * s, v2.z * c - k.z * s);
  }
  void rebake() {
    tx = f3(1.0f, 0.0f, 0.0f);
    ty = f3(0.0f, 1.0f, 0.0f);
    tz = f3(0.0f, 0.0f, 1.0f);
    const F3 r = scale(euler, 1.0f / kRadian);
    turn_pair(ty, tx, tz, -r.y);
    turn_pair(tx, ty, tz, -r.x);
    turn_pair(tz, tx, ty, -r.z);
    tx = hadamard(tx, scl);
    ty = hadamard(ty, scl);
    tz = hadamard(tz, scl);
  }
  F3 point(F3 v) const { return f3(dot(v, tx) + pos.x, dot(v, ty) + pos.y, dot(v, tz) + pos.z); }
  // BakedTriFace >> Transform (BakedTriMesh.impl.h:16-18): normals go through the point map
  // and lose the translation again.
  F3 normal(F3 n) const { return sub(point(n), pos); }
  F3 right() const { return divide(tx, scl); }
  F3 up() const { return divide(ty, scl); }
  F3 front() const { return divide(tz, scl); }
};

// ------------------------------------------------------------------ Dson
// Dson.cu:224-447: JSON-like text.  Tokens are split by whitespace and by the
// characters { } [ ] , : ; strings are quoted with the escapes \" \\ \/ \b \f \n
// \r \t (any other backslash stays literal); true / false / null are literals;
// any other bare token is a number read the way `stream >> double` reads it (a
// numeric prefix, the rest of the token ignored).  Dicts reject duplicate keys.
struct DValue {
  enum Kind { DICT, ARRAY, STRING, NUMBER, BOOL, NUL } kind = NUL;
  std::map<std::string, std::unique_ptr<DValue>> dict;
  std::vector<std::unique_ptr<DValue>> arr;
  std::string str;
  double num = 0.0;
  bool b = false;

  bool has(const char *k) const { return kind == DICT && dict.count(k) != 0; }
  const DValue &at(const char *k) const { return *dict.at(k); }
  float f() const { return (float)num; }      // Number::floatValue
  int i() const { return (int)num; }          // Number::intValue
};

class DsonParser {
 public:
  explicit DsonParser(const std::string &t) : s_(t) {}
  std::unique_ptr<DValue> parse(std::string &err) {
    skip();
    if (p_ >= s_.size()) { err = "Dson: no value in the text"; return nullptr; }
    std::unique_ptr<DValue> v = value(err);
    if (!v) return nullptr;
    skip();
    if (p_ < s_.size()) { err = "Dson: more than one value at the top level"; return nullptr; }
    return v;
  }

 private:
  static bool space(char c) { return c == ' ' || (c >= '\t' && c <= '\r'); }
  static bool structural(char c) { return c == '{' || c == '}' || c == '[' || c == ']' || c == ',' || c == ':'; }
  void skip() { while (p_ < s_.size() && space(s_[p_])) p_++; }

  std::unique_ptr<DValue> value(std::string &err) {
    skip();
    if (p_ >= s_.size()) { err = "Dson: unexpected end of text"; return nullptr; }
    const char c = s_[p_];
    if (c == '{') return dict(err);
    if (c == '[') return array(err);
    if (c == '"') return string(err);
    if (structural(c)) { err = std::string("Dson: unexpected '") + c + "'"; return nullptr; }
    size_t e = p_;
    while (e < s_.size() && !space(s_[e]) && !structural(s_[e]) && s_[e] != '"') e++;
    const std::string tok = s_.substr(p_, e - p_);
    p_ = e;
    auto v = std::make_unique<DValue>();
    if (tok == "true") { v->kind = DValue::BOOL; v->b = true; }
    else if (tok == "false") { v->kind = DValue::BOOL; v->b = false; }
    else if (tok == "null") { v->kind = DValue::NUL; }
    else if (number(tok, v->num)) { v->kind = DValue::NUMBER; }
    else { err = "Could not categorize token '" + tok + "'"; return nullptr; }
    return v;
  }

  // std::num_get's float grammar: sign, digits, '.', digits, then an exponent only
  // after mantissa digits; the collected characters must all convert (strtod).
  static bool number(const std::string &t, double &out) {
    std::string acc;
    size_t i = 0;
    bool digits = false;
    if (i < t.size() && (t[i] == '+' || t[i] == '-')) acc += t[i++];
    while (i < t.size() && std::isdigit((unsigned char)t[i])) { acc += t[i++]; digits = true; }
    if (i < t.size() && t[i] == '.') {
      acc += t[i++];
      while (i < t.size() && std::isdigit((unsigned char)t[i])) { acc += t[i++]; digits = true; }
    }
    if (!digits) return false;
    if (i < t.size() && (t[i] == 'e' || t[i] == 'E')) {
      acc += t[i++];
      if (i < t.size() && (t[i] == '+' || t[i] == '-')) acc += t[i++];
      while (i < t.size() && std::isdigit((unsigned char)t[i])) acc += t[i++];
    }
    char *end = nullptr;
    out = std::strtod(acc.c_str(), &end);
    return end != nullptr && *end == '\0';
  }

  std::unique_ptr<DValue> string(std::string &err) {
    p_++;   // opening quote
    auto v = std::make_unique<DValue>();
    v->kind = DValue::STRING;
    while (p_ < s_.size() && s_[p_] != '"') {
      if (s_[p_] == '\\' && p_ + 1 < s_.size()) {
        const char n = s_[p_ + 1];
        const char *from = "\"\\/bfnrt", *to = "\"\\/\b\f\n\r\t";
        const char *hit = std::strchr(from, n);
        if (n != '\0' && hit) { v->str += to[hit - from]; p_ += 2; continue; }
      }
      v->str += s_[p_++];
    }
    if (p_ >= s_.size()) { err = "Dson: unterminated string"; return nullptr; }
    p_++;   // closing quote
    return v;
  }

  std::unique_ptr<DValue> dict(std::string &err) {
    p_++;
    auto v = std::make_unique<DValue>();
    v->kind = DValue::DICT;
    skip();
    if (p_ < s_.size() && s_[p_] == '}') { p_++; return v; }
    while (true) {
      skip();
      if (p_ >= s_.size() || s_[p_] != '"') { err = "Dict only accepts strings as keys"; return nullptr; }
      std::unique_ptr<DValue> key = string(err);
      if (!key) return nullptr;
      skip();
      if (p_ >= s_.size() || s_[p_] != ':') { err = "Expected ':' after a key"; return nullptr; }
      p_++;
      std::unique_ptr<DValue> item = value(err);
      if (!item) return nullptr;
      if (v->dict.count(key->str)) { err = "Dict has a duplicate entry for '" + key->str + "'"; return nullptr; }
      v->dict[key->str] = std::move(item);
      skip();
      if (p_ < s_.size() && s_[p_] == ',') { p_++; continue; }
      if (p_ < s_.size() && s_[p_] == '}') { p_++; return v; }
      err = "Expected ',' after a key : value pair";
      return nullptr;
    }
  }

  std::unique_ptr<DValue> array(std::string &err) {
    p_++;
    auto v = std::make_unique<DValue>();
    v->kind = DValue::ARRAY;
    skip();
    if (p_ < s_.size() && s_[p_] == ']') { p_++; return v; }
    while (true) {
      std::unique_ptr<DValue> item = value(err);
      if (!item) return nullptr;
      v->arr.push_back(std::move(item));
      skip();
      if (p_ < s_.size() && s_[p_] == ',') { p_++; continue; }
      if (p_ < s_.size() && s_[p_] == ']') { p_++; return v; }
      err = "Array elements should be separated by ',' character";
      return nullptr;
    }
  }

  const std::string &s_;
  size_t p_ = 0;
};

// Vector3::fromDson (Vector3.impl.h:750-767): the first three numbers of an array
// (other elements skipped), missing components 0.
bool vec_from(const DValue &v, F3 &out, std::string &err) {
  if (v.kind != DValue::ARRAY) { err = "Vector3 can not be constructed from any Dson::Object other than Dson::Array"; return false; }
  float c[3] = {0.0f, 0.0f, 0.0f};
  int k = 0;
  for (const auto &e : v.arr)
    if (e->kind == DValue::NUMBER) {
      if (k < 3) c[k] = e->f();
      k++;
    }
  out = f3(c[0], c[1], c[2]);
  return true;
}

// Transform::fromDson (Transform.impl.h:209-237)
bool transform_from(const DValue &v, Transform &out, std::string &err) {
  if (v.kind != DValue::DICT) { err = "Transform can not be constructed from any Dson::Object other than Dson::Dict"; return false; }
  Transform t;
  F3 vec{};
  if (v.has("position")) {
    if (!vec_from(v.at("position"), vec, err)) return false;
    t.pos = vec;
  }
  if (v.has("rotation")) {
    if (!vec_from(v.at("rotation"), vec, err)) return false;
    t.euler = vec;
    t.rebake();
  }
  if (v.has("scale")) {
    if (!vec_from(v.at("scale"), vec, err)) return false;
    t.scl = vec;
    t.rebake();
  }
  if (v.has("relative_position")) {
    if (!vec_from(v.at("relative_position"), vec, err)) return false;
    const F3 d = add(add(scale(t.right(), vec.x), scale(t.up(), vec.y)), scale(t.front(), vec.z));
    t.pos = add(t.pos, d);
  }
  out = t;
  return true;
}

// ------------------------------------------------------------------ OBJ
// MeshReader.impl.h.  Positions and normals are mirrored in x (:182-184); `g`
// starts a named object (`o` is ignored); faces index 1-based v/vt/vn.  Vertex
// normals are the normalised sums of the corner cross products over every face of
// the file (:202-229); a face corner without a normal index takes, through the
// per-object index map, the vertex normal of the first corner of that object that
// had the same (missing) index (:313-323).
struct Corner {
  int v, t, n;
};
struct ObjMesh {
  std::vector<std::vector<F3>> face_pos, face_nrm, face_tex;   // per face, per corner
};

class ObjReader {
 public:
  explicit ObjReader(const std::string &text) : f_(text.c_str()) {}

  bool read(std::vector<ObjMesh> &meshes, std::vector<std::string> &names, std::string &err) {
    std::vector<F3> verts, norms, texs;
    std::vector<std::vector<Corner>> faces;
    std::vector<std::pair<int, int>> objects;
    int start = 0;
    std::string tok;
    while (f_[c_] != '\0') {
      if (!token(tok, "", "", "#")) continue;
      if (tok == "v") verts.push_back(vertex(-1.0f));
      else if (tok == "vt") texs.push_back(vertex(1.0f));
      else if (tok == "vn") norms.push_back(vertex(-1.0f));
      else if (tok == "f") {
        faces.emplace_back();
        std::string nt;
        while (token(nt, "", "\n", "#")) faces.back().push_back(corner(nt));
      } else if (tok == "g") {
        if ((int)faces.size() > start) {
          objects.emplace_back(start, (int)faces.size());
          start = (int)faces.size();
        } else {
          names.clear();
        }
        names.emplace_back();
        line(names.back());
      }
    }
    if ((int)faces.size() > start) objects.emplace_back(start, (int)faces.size());
    else names.clear();
    if (objects.empty()) return true;

    std::vector<F3> vn(verts.size(), f3(0.0f, 0.0f, 0.0f));
    const int nv = (int)verts.size();
    for (const auto &fc : faces) {
      const int n = (int)fc.size();
      for (int j = 0; j < n; j++) {
        const int pv = fc[(j - 1 + n) % n].v, nx = fc[(j + 1) % n].v, cv = fc[j].v;
        if (pv < 0 || pv >= nv || nx < 0 || nx >= nv || cv < 0 || cv >= nv) {
          err = "face with a vertex index out of range";
          return false;
        }
        vn[cv] = add(vn[cv], cross(sub(verts[pv], verts[cv]), sub(verts[nx], verts[cv])));
      }
    }
    for (auto &n : vn)
      if (n.x != 0.0f || n.y != 0.0f || n.z != 0.0f) n = normalized(n);

    for (const auto &ob : objects) {
      ObjMesh m;
      std::unordered_map<int, F3> nmap;
      for (int i = ob.first; i < ob.second; i++) {
        std::vector<F3> P, N, T;
        for (const Corner &k : faces[i]) {
          if (k.v < 0 || k.v >= nv) { err = "vertex index out of range"; return false; }
          auto it = nmap.find(k.n);
          if (it == nmap.end()) {
            if (k.n >= (int)norms.size()) { err = "normal index out of range"; return false; }
            it = nmap.emplace(k.n, k.n >= 0 ? norms[k.n] : vn[k.v]).first;
          }
          if (k.t >= (int)texs.size()) { err = "texture index out of range"; return false; }
          P.push_back(verts[k.v]);
          N.push_back(it->second);
          T.push_back(k.t >= 0 ? texs[k.t] : f3(0.0f, 0.0f, 0.0f));   // a corner without one: Vertex::zero()
        }
        m.face_pos.push_back(std::move(P));
        m.face_nrm.push_back(std::move(N));
        m.face_tex.push_back(std::move(T));
      }
      meshes.push_back(std::move(m));
    }
    return true;
  }

 private:
  static bool space(char c) { return c == ' ' || (c >= '\t' && c <= '\r'); }
  static bool in(const char *set, char c) { return c != '\0' && std::strchr(set, c) != nullptr; }

  // moveToRelevant (:18-24)
  bool relevant(const char *bp, const char *instant, const char *lb) {
    while (f_[c_] != '\0' && (space(f_[c_]) || in(bp, f_[c_]))) {
      if (in(instant, f_[c_])) return false;
      c_++;
    }
    return !in(lb, f_[c_]);
  }
  // getLine (:26-36): skips whitespace (newlines too), takes the rest of the line,
  // trailing whitespace trimmed.
  bool line(std::string &dst) {
    dst.clear();
    if (relevant("", "", ""))
      while (f_[c_] != '\n' && f_[c_] != '\0') dst += f_[c_++];
    while (!dst.empty() && space(dst.back())) dst.pop_back();
    return !dst.empty();
  }
  // getToken (:38-52)
  bool token(std::string &dst, const char *bp, const char *instant, const char *lb) {
    dst.clear();
    if (relevant(bp, instant, lb)) {
      while (f_[c_] != '\0' && !(space(f_[c_]) || in(bp, f_[c_]))) dst += f_[c_++];
    } else if (f_[c_] != '\0' && in(lb, f_[c_])) {
      std::string skipped;
      line(skipped);
      return false;
    }
    return !dst.empty();
  }
  // getFloat from this line (:60-65): atof, 0 when the line has no more tokens
  float number() {
    std::string t;
    if (!token(t, "", "\n", "#")) return 0.0f;
    return (float)std::atof(t.c_str());
  }
  F3 vertex(float sx) {
    const float x = number(), y = number(), z = number();
    return hadamard(f3(x, y, z), f3(sx, 1.0f, 1.0f));
  }
  // getInt / getNode (:81-98): atoi of the text up to each '/', 0 when empty, minus 1
  static Corner corner(const std::string &t) {
    size_t p = 0;
    int vals[3];
    for (int k = 0; k < 3; k++) {
      std::string d;
      while (p < t.size() && t[p] != '/') d += t[p++];
      if (p == t.size()) p++;   // the field ran to the end of the token: later fields are empty
      vals[k] = (d.empty() ? 0 : std::atoi(d.c_str())) - 1;
      p++;
    }
    return Corner{vals[0], vals[1], vals[2]};
  }

  const char *f_;
  size_t c_ = 0;
};

// ------------------------------------------------------------------ scene context
struct Unit {
  F3 v[3], n[3], t[3];
  int mat;
};

struct Item {
  std::string file, object;
  Transform t;
  int mat = -1;
};
struct SubGroup {
  std::string group;
  Transform t;
  int fallback = -1, override_ = -1;
};
struct Group {
  std::string name;
  std::vector<Item> items;
  std::vector<SubGroup> subs;
  Transform t;
  int fallback = -1, override_ = -1;
};

// DumbBasedShader state (DumbBasedShader.cuh:25-33) and its constructor (.impl.cuh:5-16)
struct DumbBased {
  F3 fresnel, diffuse;
  float spec, metal;
  static DumbBased make(F3 fresnel, float spec, F3 diffuse, float metal) {
    DumbBased d;
    d.fresnel = fresnel;
    const float m = std::max(std::min(metal, 1.0f), 0.0f);
    d.diffuse = scale(diffuse, 1.0f - m);
    d.spec = spec;
    d.metal = m;
    return d;
  }
};

}  // namespace

struct dr_scene_file_s {
  std::string source_path;
  std::unordered_map<std::string, int> material_names;
  std::unordered_map<std::string, std::unordered_map<std::string, ObjMesh>> obj_files;
  std::map<std::string, Group> groups;
  std::vector<dr_material> materials;
  std::vector<dr_material_maps> maps;
  std::vector<dr_texture> textures;
  std::vector<float> texels;
  std::unordered_map<std::string, int> texture_names;
  std::vector<dr_light> lights;
  std::vector<Unit> geometry;
  Transform camera;
  bool lens_set = false;
  float lens_x = 0.0f;
  dr_scene_file_settings settings{};
  std::string err;

  dr_scene_file_s() {
    // DumbRenderContext::DumbRenderContext (DumbRenderContext.cu:282-289)
    settings.boxing = DR_BOXING_HEIGHT;
    settings.max_bounces = 2;
    settings.fsaa_x = settings.fsaa_y = 1;
    settings.ignore_backfaces = 1;
    settings.lens = DR_STOCHASTIC_LENS;
  }

  bool error(const std::string &m) {
    if (err.empty()) err = m;
    return false;
  }

  // SimpleStochasticLense(angle) (SimpleStochasticLense.impl.cuh:4-9)
  static float lens_focal(float angle) {
    if (angle < 1) angle = 1;
    else if (angle > 179) angle = 179;
    const float radian = angle / (2 * kRadian);
    return 0.5f / std::tan(radian);
  }

  static bool open(const std::string &path) {
    std::ifstream f(path);
    return !f.fail();
  }

  // ---------------------------------------------------------------- files
  bool from_file(const std::string &filename, int depth) {
    if (depth > 32) return error("Error: include depth exceeds 32 (recursive includes?)");
    std::ifstream stream(filename);
    if (stream.fail()) return error("Error: Could not open file: \"" + filename + "\"..");
    const std::string text((std::istreambuf_iterator<char>(stream)), std::istreambuf_iterator<char>());
    size_t len = filename.size();
    while (len > 0 && filename[len - 1] != '/' && filename[len - 1] != '\\') len--;
    return from_text(text, filename.substr(0, len), depth);
  }
  bool from_text(const std::string &text, const std::string &dir, int depth) {
    std::string perr;
    std::unique_ptr<DValue> root = DsonParser(text).parse(perr);
    if (!root) return error("Error: Could not parse scene text: " + perr);
    const std::string saved = source_path;
    source_path = dir;
    const bool ok = from_dson(*root, depth);
    source_path = saved;
    return ok;
  }

  // DumbRenderContext::fromDson (:337-371)
  bool from_dson(const DValue &d, int depth) {
    if (d.kind != DValue::DICT) return error("Error: Render context can be constructed only from a Dson::Dict");
    if (d.has("include") && !include(d.at("include"), depth)) return false;
    if (d.has("materials")) {
      const DValue &a = d.at("materials");
      if (a.kind != DValue::ARRAY) return error("Error: Materials should be contained in Dson::Array");
      for (const auto &m : a.arr)
        if (!parse_material(*m, nullptr)) return false;
    }
    if (d.has("lights")) {
      const DValue &a = d.at("lights");
      if (a.kind != DValue::ARRAY) return error("Error: Lights should be contained in Dson::Array");
      for (const auto &l : a.arr)
        if (!parse_light(*l)) return false;
    }
    if (d.has("groups")) {
      const DValue &a = d.at("groups");
      if (a.kind != DValue::ARRAY) return error("Error: Groups should be contained in Dson::Array");
      for (const auto &g : a.arr) {
        Group grp;
        if (!group_from(*g, grp)) return false;
      }
    }
    if (d.has("objects")) {
      const DValue &a = d.at("objects");
      if (a.kind != DValue::ARRAY) return error("Error: Objects should be contained in Dson::Array");
      for (const auto &o : a.arr)
        if (!parse_object(*o)) return false;
    }
    if (d.has("camera") && !parse_camera(d.at("camera"))) return false;
    if (!lens_set) {
      lens_set = true;
      lens_x = lens_focal(64.0f);
    }
    if (d.has("renderer") && !parse_renderer(d.at("renderer"))) return false;
    return true;
  }

  // includeFiles (:641-659): relative to the including file when that path opens
  bool include(const DValue &v, int depth) {
    if (v.kind != DValue::ARRAY) return error("Error: Include should be a list of strings");
    for (const auto &e : v.arr) {
      if (e->kind != DValue::STRING) return error("Error: Can not include non-string objects");
      std::string path = source_path + e->str;
      if (!open(path)) path = e->str;
      if (!from_file(path, depth + 1)) return false;
    }
    return true;
  }

  // loadObjFile (:727-752): cached under the name as written
  bool load_obj(const std::string &filename) {
    const std::string key = "obj::" + filename;
    if (obj_files.count(key)) return true;
    std::string path = source_path + filename;
    if (!open(path)) path = filename;
    std::ifstream stream(path);
    if (stream.fail()) return error("Error: Could not read file: '" + filename + "' (" + path + ")");
    const std::string text((std::istreambuf_iterator<char>(stream)), std::istreambuf_iterator<char>());
    std::vector<ObjMesh> meshes;
    std::vector<std::string> names;
    std::string rerr;
    if (!ObjReader(text).read(meshes, names, rerr))
      return error("Error: Could not read file: '" + filename + "' (" + path + "): " + rerr);
    if (meshes.size() != names.size())
      return error("Error: File '" + filename + "' does not have equal amount of names and meshes");
    auto &dict = obj_files[key];
    for (size_t i = 0; i < meshes.size(); i++) dict[names[i]] = std::move(meshes[i]);
    return true;
  }

  // ---------------------------------------------------------------- textures
  // DumbRenderContext::getImageId (DumbRenderContext.cu:374-452): bool / null / number = no
  // texture; a string names a texture ("name::" first, then "png::"); a dict loads its "png"
  // file (next to the scene file when that path opens) as a new texture, with optional
  // "filtering" (none / [bilinear]) and "name".
  bool image_id(const DValue &v, int &id) {
    if (v.kind == DValue::BOOL || v.kind == DValue::NUL || v.kind == DValue::NUMBER) { id = -1; return true; }
    if (v.kind == DValue::STRING) {
      auto it = texture_names.find("name::" + v.str);
      if (it == texture_names.end()) it = texture_names.find("png::" + v.str);
      if (it == texture_names.end()) return error("Error: Texture \"" + v.str + "\" not found");
      id = it->second;
      return true;
    }
    if (v.kind != DValue::DICT) return error("Error: Unsupported dson type for texture");
    if (!v.has("png")) return error("Error: Image dict incomplete");
    if (v.at("png").kind != DValue::STRING) return error("Error: Image 'png' entry MUST BE a string");
    const std::string &name = v.at("png").str;
    std::string path = source_path + name;
    if (!open(path)) path = name;
    std::vector<unsigned char> rgba;
    uint32_t w = 0, h = 0;
    std::string perr;
    if (!dr::png_decode_file(path, rgba, w, h, perr)) return error("Error: Could not read file: \"" + path + "\"");
    id = (int)textures.size();
    dr_texture t{};
    t.offset = (int32_t)(texels.size() / 4);
    t.width = (int32_t)w;
    t.height = (int32_t)h;
    t.filter = 1;                                   // Texture's default FILTER_BILINEAR
    for (unsigned char b : rgba) texels.push_back(((float)b) / 255.0f);   // Images::getTexturePNG
    textures.push_back(t);
    texture_names["png::" + path] = id;
    texture_names["png::" + name] = id;
    if (v.has("filtering")) {
      const DValue &f = v.at("filtering");
      if (f.kind != DValue::STRING) return error("Error: Image 'filtering' entry MUST BE a string");
      if (f.str == "none") textures[id].filter = 0;
      else if (f.str == "bilinear") textures[id].filter = 1;
      else return error("Error: Image filter can be only \"none\"/[\"bilinear\"] (got: \"" + f.str + "\")");
    }
    if (v.has("name")) {
      if (v.at("name").kind != DValue::STRING) return error("Error: Image name entry MUST BE a string");
      texture_names["name::" + v.at("name").str] = id;
    }
    return true;
  }

  // ---------------------------------------------------------------- plugins
  // ColoredTexture::fromDson (DumbStructs.cu:7-37): colour, texture, tiling and offset keys.
  bool colored(const DValue &d, const char *ck, const char *tk, const char *tl, const char *of, F3 &color,
               dr_texture_map &map) {
    if (d.has(ck) && !vec_from(d.at(ck), color, err)) return false;
    if (d.has(tk) && !image_id(d.at(tk), map.texture)) return false;
    F3 v{};
    if (d.has(tl)) {
      if (!vec_from(d.at(tl), v, err)) return false;
      map.tiling[0] = v.x; map.tiling[1] = v.y;
    }
    if (d.has(of)) {
      if (!vec_from(d.at(of), v, err)) return false;
      map.offset[0] = v.x; map.offset[1] = v.y;
    }
    return true;
  }
  static dr_texture_map no_texture() {
    dr_texture_map m{};
    m.texture = -1;
    m.tiling[0] = m.tiling[1] = 1.0f;
    return m;
  }
  // Keeps the material's maps when any of them has a texture (dr_material.reserved[0]).
  void attach_maps(dr_material &out, const dr_material_maps &mp) {
    bool any = false;
    for (int k = 0; k < 4; k++) any = any || mp.map[k].texture >= 0;
    if (!any) return;
    maps.push_back(mp);
    out.reserved[0] = (int32_t)maps.size();
  }
  bool number_of(const DValue &d, const char *k, float &out, const char *msg) {
    if (!d.has(k)) return true;
    if (d.at(k).kind != DValue::NUMBER) return error(msg);
    out = d.at(k).f();
    return true;
  }

  // DumbBasedShader::fromDson (DumbBasedShader.impl.cuh:125-170)
  bool dumb_based(const DValue &d, dr_material &out) {
    DumbBased s = DumbBased::make(f3(0.2f, 0.2f, 0.2f), 8.0f, f3(0.2f, 0.2f, 0.2f), 0.5f);
    if (d.has("preset")) {
      const DValue &p = d.at("preset");
      if (p.kind != DValue::STRING) return error("Error: DumbBasedShader preset should be of a string type...");
      // DumbBasedShader.cuh:14-17
      if (p.str == "rough_gold") s = DumbBased::make(f3(1.022f, 0.782f, 0.344f), 16.0f, f3(1.0f, 1.0f, 1.0f), 1.0f);
      else if (p.str == "glossy_gold" || p.str == "gold")
        s = DumbBased::make(f3(1.022f, 0.782f, 0.344f), 128.0f, f3(1.0f, 1.0f, 1.0f), 1.0f);
      else if (p.str == "glossy_finish" || p.str == "glossy")
        s = DumbBased::make(f3(0.175f, 0.15f, 0.125f), 4096.0f, f3(0.32f, 0.32f, 0.32f), 0.25f);
      else if (p.str == "matte_finish" || p.str == "matte")
        s = DumbBased::make(f3(0.175f, 0.15f, 0.125f), 8.0f, f3(0.24f, 0.24f, 0.24f), 0.05f);
      else return error("Error: DumbBasedShader preset \"" + p.str + "\" does not exist...");
    }
    F3 fresnel = s.fresnel;
    float metal = s.metal, specular = s.spec;
    F3 diffuse = scale(s.diffuse, 1.0f / ((metal == 1.0f) ? 1.0f : (1.0f - metal)));   // Color /= f
    F3 normal_color = f3(1.0f, 1.0f, 1.0f), alpha = f3(1.0f, 1.0f, 1.0f);
    dr_material_maps mp{};
    for (int k = 0; k < 4; k++) mp.map[k] = no_texture();
    if (!colored(d, "fresnel", "fresnel_texture", "fresnel_tiling", "fresnel_offset", fresnel, mp.map[0])) return false;
    if (!colored(d, "diffuse", "diffuse_texture", "diffuse_tiling", "diffuse_offset", diffuse, mp.map[1])) return false;
    if (!colored(d, "normal_color", "normal_texture", "normal_tiling", "normal_offset", normal_color, mp.map[2]))
      return false;
    if (!colored(d, "alpha_cutout_color", "alpha_cutout_texture", "alpha_cutout_tiling", "alpha_cutout_offset", alpha,
                 mp.map[3]))
      return false;
    float thresh = 0.5f;
    if (!number_of(d, "alpha_cutout_threshold", thresh, "Error: DumbBasedShader alpha_cutout_threshold value hat to be a number..."))
      return false;
    if (!number_of(d, "specular", specular, "Error: DumbBasedShader specular value hat to be a number...")) return false;
    if (!number_of(d, "metal", metal, "Error: DumbBasedShader metal value hat to be a number...")) return false;
    s = DumbBased::make(fresnel, specular, diffuse, metal);
    out.kind = DR_DUMB_BASED;
    out.albedo[0] = s.fresnel.x; out.albedo[1] = s.fresnel.y; out.albedo[2] = s.fresnel.z;
    out.diffuse[0] = s.diffuse.x; out.diffuse[1] = s.diffuse.y; out.diffuse[2] = s.diffuse.z;
    out.p[0] = s.spec; out.p[1] = s.metal;
    mp.normal_color[0] = normal_color.x; mp.normal_color[1] = normal_color.y; mp.normal_color[2] = normal_color.z;
    mp.alpha_threshold = thresh;
    attach_maps(out, mp);
    return true;
  }

  // SimpleStochasticShader constructor (SimpleStochasticShader.impl.cuh:4-11)
  static void simple_stochastic_make(F3 color, float diffuse, float smoothness, float shininess, dr_material &out) {
    const float total = diffuse + smoothness;
    const float mul = (total > 1.0f) ? (1.0f / total) : 1.0f;
    out.kind = DR_SIMPLE_STOCHASTIC;
    out.albedo[0] = color.x; out.albedo[1] = color.y; out.albedo[2] = color.z;
    out.p[0] = diffuse * mul; out.p[1] = smoothness * mul; out.p[2] = shininess;
  }
  // SimpleStochasticShader::fromDson (:40-65), starting from the default shader
  bool simple_stochastic(const DValue &d, dr_material &out) {
    dr_material def{};
    simple_stochastic_make(f3(1.0f, 1.0f, 1.0f), 0.05f, 0.95f, 2.0f, def);
    F3 color = f3(def.albedo[0], def.albedo[1], def.albedo[2]);
    float diffuse = def.p[0], smooth = def.p[1], shine = def.p[2];
    if (d.has("color") && !vec_from(d.at("color"), color, err)) return false;
    if (!number_of(d, "diffuse", diffuse, "Error: SimpleStochasticShader diffuse has to be a number.")) return false;
    if (!number_of(d, "smoothness", smooth, "Error: SimpleStochasticShader smoothness has to be a number.")) return false;
    if (!number_of(d, "shininess", shine, "Error: SimpleStochasticShader shininess has to be a number.")) return false;
    simple_stochastic_make(color, diffuse, smooth, shine, out);
    return true;
  }

  // Shaders::FresnelShader::fromDson (FresnelShader.impl.cuh:21-25; engine colours
  // DumbStructs.cuh:143-147, an unknown name or a non-string is reported but ignored)
  bool fresnel_color(const DValue &d, dr_material &out) {
    F3 color = f3(1.0f, 1.0f, 1.0f);
    if (d.has("engine_color") && d.at("engine_color").kind == DValue::STRING) {
      const std::string &c = d.at("engine_color").str;
      if (c == "iron") color = f3(0.56f, 0.57f, 0.58f);
      else if (c == "copper") color = f3(0.95f, 0.64f, 0.54f);
      else if (c == "gold") color = f3(1.0f, 0.71f, 0.29f);
      else if (c == "aluminum") color = f3(0.91f, 0.92f, 0.92f);
      else if (c == "silver") color = f3(0.95f, 0.93f, 0.88f);
    }
    dr_material_maps mp{};
    for (int k = 0; k < 4; k++) mp.map[k] = no_texture();
    if (!colored(d, "color", "texture", "tiling", "offset", color, mp.map[0])) return false;
    out.kind = DR_FRESNEL;
    out.albedo[0] = color.x; out.albedo[1] = color.y; out.albedo[2] = color.z;
    mp.normal_color[0] = mp.normal_color[1] = mp.normal_color[2] = 1.0f;
    mp.alpha_threshold = 0.5f;
    attach_maps(out, mp);
    return true;
  }

  // parseMaterial (:662-698)
  bool parse_material(const DValue &d, int *id) {
    if (d.kind != DValue::DICT) return error("Error: Material should be contained in Dson::Dict");
    if (!d.has("type")) return error("Error: Material has to have a type");
    const DValue &t = d.at("type");
    if (t.kind != DValue::STRING) return error("Error: Material type has to be a string");
    dr_material m{};
    bool ok;
    if (t.str == "dumb_based") ok = dumb_based(d, m);
    else if (t.str == "simple_stochastic") ok = simple_stochastic(d, m);
    else if (t.str == "fresnel_color") ok = fresnel_color(d, m);
    else return error("Error: Unknown material type: \"" + t.str + "\"");
    if (!ok) return false;
    if (id) *id = (int)materials.size();
    materials.push_back(m);
    if (d.has("name") && d.at("name").kind == DValue::STRING) material_names[d.at("name").str] = (int)materials.size() - 1;
    return true;
  }

  // getMaterialId (:771-784)
  bool material_id(const DValue &d, int &id) {
    if (d.kind == DValue::STRING) {
      auto it = material_names.find(d.str);
      if (it == material_names.end()) return error("Error: Material not found: \"" + d.str + "\"");
      id = it->second;
      return true;
    }
    return parse_material(d, &id);
  }

  // parseLight (:699-725) with SimpleSoftDirectionalLight::fromDson
  // (SimpleSoftDirectionalLight.impl.cuh:22-49) and SphericalLightEmitter::fromDson
  // (SphericalLightEmitter.impl.cuh:21-41), each starting from its default light.
  bool parse_light(const DValue &d) {
    if (d.kind != DValue::DICT) return error("Error: Light should be contained in Dson::Dict");
    if (!d.has("type")) return error("Error: Light has to have a type");
    const DValue &t = d.at("type");
    if (t.kind != DValue::STRING) return error("Error: Light type has to be a string");
    dr_light l{};
    if (t.str == "simple_soft_directional") {
      // the default light (SimpleSoftDirectionalLight.cuh:16: direction (1,-1,1).normalized(),
      // distance 1024, softness 0.125) as its constructor stores it (.impl.cuh:5-10)
      F3 shade = f3(1.0f, 1.0f, 1.0f);
      F3 dir = normalized(normalized(f3(1.0f, -1.0f, 1.0f)));
      const float dist0 = 1024.0f, soft0 = dist0 * 0.125f;
      float distance = dist0;
      float softness = soft0 / ((distance == 0.0f) ? 0.0f : distance);
      if (d.has("color") && !vec_from(d.at("color"), shade, err)) return false;
      if (d.has("direction")) {
        dir = f3(0.0f, 0.0f, 0.0f);
        if (!vec_from(d.at("direction"), dir, err)) return false;
      }
      if (!number_of(d, "distance", distance, "Error: SimpleSoftDirectionalLight distance has to have a numeric value..."))
        return false;
      if (!number_of(d, "softness", softness, "Error: SimpleSoftDirectionalLight softness has to have a numeric value..."))
        return false;
      const F3 u = normalized(dir);
      l.kind = DR_SOFT_DIRECTIONAL_LIGHT;
      l.emission[0] = shade.x; l.emission[1] = shade.y; l.emission[2] = shade.z;
      l.edge_u[0] = u.x; l.edge_u[1] = u.y; l.edge_u[2] = u.z;
      l.dist = distance;
      l.radius = distance * softness;
    } else if (t.str == "simple_spherical") {
      F3 shade = f3(1.0f, 1.0f, 1.0f), pos = f3(0.0f, 0.0f, 0.0f);
      float radius = 1.0f;
      if (d.has("color") && !vec_from(d.at("color"), shade, err)) return false;
      if (d.has("position") && !vec_from(d.at("position"), pos, err)) return false;
      if (!number_of(d, "radius", radius, "Error: SphericalLightEmitter radius has to be a number.")) return false;
      l.kind = DR_SPHERICAL_LIGHT;
      l.emission[0] = shade.x; l.emission[1] = shade.y; l.emission[2] = shade.z;
      l.corner[0] = pos.x; l.corner[1] = pos.y; l.corner[2] = pos.z;
      l.radius = radius;
    } else if (t.str == "spotlight") {
      if (!spotlight(d, l)) return false;
    } else {
      return error("Error: Unknown light type: \"" + t.str + "\"");
    }
    lights.push_back(l);
    return true;
  }

  // Lights::Spotlight::fromDson (Spotlight.impl.cuh:95-161) and its constructor (:7-38)
  bool spotlight(const DValue &d, dr_light &l) {
    F3 shade = f3(1.0f, 1.0f, 1.0f), pos = f3(0.0f, 0.0f, 0.0f), dir = f3(0.0f, 0.0f, 1.0f);
    float lum = 1500.0f, inner = 16.0f, outer = 64.0f, falloff = 1.0f, size = 0.25f;
    bool sphere = false, shadows = true;
    int samples = 1;
    if (d.has("color") && !vec_from(d.at("color"), shade, err)) return false;
    if (!number_of(d, "luminosity", lum, "Error: Spotlight luminosity must be a number")) return false;
    if (d.has("position") && !vec_from(d.at("position"), pos, err)) return false;
    if (d.has("target")) {
      F3 target{};
      if (!vec_from(d.at("target"), target, err)) return false;
      dir = sub(target, pos);
    } else if (d.has("direction") && !vec_from(d.at("direction"), dir, err)) {
      return false;
    }
    if (!number_of(d, "hotspot_angle", inner, "Error: Spotlight hotspot angle must be a number")) return false;
    if (!number_of(d, "falloff_angle", outer, "Error: Spotlight falloff angle must be a number")) return false;
    if (!number_of(d, "falloff_power", falloff, "Error: Spotlight falloff power must be a number")) return false;
    if (!number_of(d, "emitter_size", size, "Error: Spotlight emitter size must be a number")) return false;
    if (d.has("emitter_shape")) {
      const DValue &e = d.at("emitter_shape");
      if (e.kind != DValue::STRING) return error("Error: Spotlight emitter shape must be a string");
      if (e.str == "disc") sphere = false;
      else if (e.str == "sphere") sphere = true;     // any other name is reported and ignored
    }
    if (d.has("cast_shadows")) {
      if (d.at("cast_shadows").kind != DValue::BOOL) return error("Error: Spotlight shadow cast flag must be a boolean");
      shadows = d.at("cast_shadows").b;
    }
    if (d.has("sample_count")) {
      if (d.at("sample_count").kind != DValue::NUMBER) return error("Error: Spotlight sample count must be a number");
      samples = d.at("sample_count").i();
    }
    // constructor
    const float kPi = 3.14159265359f, kEps = 0.000005f;
    const F3 u = normalized(dir);
    if (size < (32 * kEps)) size = (32 * kEps);
    if (inner < 0.0f) inner = (-inner);
    if (inner > 180.0f) inner = 180.0f;
    if (outer < 0.0f) outer = (-inner);
    if (outer > 180.0f) outer = 180.0f;
    if (outer < inner) outer = inner;
    l.kind = DR_SPOT_LIGHT;
    l.emission[0] = shade.x * lum; l.emission[1] = shade.y * lum; l.emission[2] = shade.z * lum;
    l.corner[0] = pos.x; l.corner[1] = pos.y; l.corner[2] = pos.z;
    l.edge_u[0] = u.x; l.edge_u[1] = u.y; l.edge_u[2] = u.z;
    l.edge_v[0] = std::cos(inner * kPi / 180.0f);
    l.edge_v[1] = std::cos(outer * kPi / 180.0f);
    l.edge_v[2] = falloff;
    l.radius = size;
    l.reserved = ((samples > 0) ? ((samples > 16) ? 16 : samples) : 0) | (shadows ? 32 : 0) | (sphere ? 64 : 0);
    return true;
  }

  // ---------------------------------------------------------------- groups and objects
  // Group::Item::fromDson (:154-191)
  bool item_from(const DValue &d, Item &it) {
    if (d.kind != DValue::DICT) return error("Error: Group item can only be made from a dictionary");
    if (!d.has("mesh")) return error("Error: Item has to have a 'mesh' entry");
    const DValue &m = d.at("mesh");
    if (m.kind != DValue::DICT) return error("Error: Item 'mesh' entry has to be a dict");
    if (m.has("obj")) {
      const DValue &o = m.at("obj");
      if (o.kind != DValue::STRING) return error("Error: .obj file name must be a string");
      if (!load_obj(o.str)) return false;
      it.file = "obj::" + o.str;
    } else if (m.has("primitive")) {
      return error("Error: primitive meshes are not implemented (the scene format reserves them)");
    } else {
      return error("Error: Object mesh could not be parsed");
    }
    if (m.has("object")) {
      const DValue &n = m.at("object");
      if (n.kind != DValue::STRING) return error("Error: .obj mesh name must be a string");
      const auto &meshes = obj_files[it.file];
      if (!meshes.count(n.str)) return error("Error: Object [" + n.str + "] not found in [" + it.file + "]");
      it.object = n.str;
    } else {
      it.object.clear();
    }
    if (d.has("transform") && !transform_from(d.at("transform"), it.t, err)) return false;
    if (d.has("material") && !material_id(d.at("material"), it.mat)) return false;
    return true;
  }

  // Group::SubGroup::fromDson (:193-219)
  bool sub_from(const DValue &d, SubGroup &sg) {
    if (d.kind == DValue::STRING) {
      sg.group = "named::" + d.str;
      if (!groups.count(sg.group)) return error("Error: Group [" + d.str + "] not found");
      return true;
    }
    if (d.kind != DValue::DICT) return error("Error: Subgroup should be expressed either as a string, or a dict");
    if (d.has("group")) {
      const DValue &g = d.at("group");
      if (g.kind != DValue::STRING)
        return error("Error: Subgroup with pre-defined group reference must have 'group' key, paired with string value");
      sg.group = "named::" + g.str;
      if (!groups.count(sg.group)) return error("Error: Group [" + sg.group + "] not found");
      if (d.has("transform") && !transform_from(d.at("transform"), sg.t, err)) return false;
      if (d.has("fallback_material") && !material_id(d.at("fallback_material"), sg.fallback)) return false;
      if (d.has("override_material") && !material_id(d.at("override_material"), sg.override_)) return false;
    } else {
      Group inner;
      if (!group_from(d, inner)) return false;
      sg.group = inner.name;
    }
    return true;
  }

  // Group::fromDson (:221-257)
  bool group_from(const DValue &d, Group &g) {
    if (d.kind != DValue::DICT) return error("Error: Group can only be made from a dictionary");
    if (d.has("elements")) {
      const DValue &el = d.at("elements");
      if (el.kind != DValue::ARRAY) return error("Error: Group elements should be contained in an array");
      for (const auto &e : el.arr) {
        if (e->kind == DValue::DICT && e->has("mesh")) {
          Item it;
          if (!item_from(*e, it)) return false;
          g.items.push_back(it);
        } else {
          SubGroup sg;
          if (!sub_from(*e, sg)) return false;
          g.subs.push_back(sg);
        }
      }
    }
    if (d.has("transform") && !transform_from(d.at("transform"), g.t, err)) return false;
    if (d.has("fallback_material") && !material_id(d.at("fallback_material"), g.fallback)) return false;
    if (d.has("override_material") && !material_id(d.at("override_material"), g.override_)) return false;
    if (d.has("name")) {
      if (d.at("name").kind != DValue::STRING) return error("Error: Group name must be a string");
      g.name = "named::" + d.at("name").str;
    } else {
      g.name = "unnamed::" + std::to_string(groups.size());
    }
    groups[g.name] = g;
    return true;
  }

  // Item::resolve (:106-123): PolyMesh::bake fans (PolyMesh.impl.h:140-152), each
  // face moved by the item's transform; an item without an object name takes every
  // mesh of the file in the file dictionary's (std::unordered_map) order.
  void resolve_mesh(const ObjMesh &m, const Item &it, std::vector<Unit> &out) const {
    for (size_t f = 0; f < m.face_pos.size(); f++) {
      const auto &P = m.face_pos[f];
      const auto &N = m.face_nrm[f];
      int prev = 1;
      for (size_t cur = 2; cur < P.size(); cur++) {
        Unit u;
        const size_t idx[3] = {0, (size_t)prev, cur};
        const auto &TC = m.face_tex[f];
        for (int k = 0; k < 3; k++) {
          u.v[k] = it.t.point(P[idx[k]]);
          u.n[k] = it.t.normal(N[idx[k]]);
          u.t[k] = TC[idx[k]];   // texture coordinates are not transformed (BakedTriMesh.impl.h:16-18)
        }
        u.mat = it.mat;
        out.push_back(u);
        prev = (int)cur;
      }
    }
  }
  void resolve_item(const Item &it, std::vector<Unit> &out) const {
    const auto &meshes = obj_files.at(it.file);
    if (!it.object.empty()) resolve_mesh(meshes.at(it.object), it, out);
    else
      for (const auto &kv : meshes) resolve_mesh(kv.second, it, out);
  }
  static void finish(std::vector<Unit> &u, size_t from, const Transform &t, int override_, int fallback) {
    for (size_t i = from; i < u.size(); i++) {
      if (override_ >= 0) u[i].mat = override_;
      else if (u[i].mat < 0) u[i].mat = fallback;
      for (int k = 0; k < 3; k++) {
        u[i].v[k] = t.point(u[i].v[k]);
        u[i].n[k] = t.normal(u[i].n[k]);
      }
    }
  }
  // Group::resolve (:131-150) and SubGroup::resolve (:124-130)
  bool resolve_group(const std::string &name, std::vector<Unit> &out, int depth) {
    if (depth > 64) return error("Error: group nesting exceeds 64 levels (a group that contains itself?)");
    const Group &g = groups.at(name);
    const size_t from = out.size();
    for (const Item &it : g.items) resolve_item(it, out);
    for (const SubGroup &sg : g.subs) {
      const size_t s0 = out.size();
      if (!resolve_group(sg.group, out, depth + 1)) return false;
      finish(out, s0, sg.t, sg.override_, sg.fallback);
    }
    finish(out, from, g.t, g.override_, g.fallback);
    return true;
  }

  // parseObject (:754-770)
  bool parse_object(const DValue &d) {
    std::vector<Unit> units;
    if (d.kind == DValue::DICT && d.has("mesh")) {
      Item it;
      if (!item_from(d, it)) return false;
      if (it.mat < 0) return error("Error: Object MUST have a material attached to it");
      resolve_item(it, units);
    } else {
      SubGroup sg;
      if (!sub_from(d, sg)) return false;
      if (!resolve_group(sg.group, units, 0)) return false;
      finish(units, 0, sg.t, sg.override_, sg.fallback);
      for (const Unit &u : units)
        if (u.mat < 0) return error("Error: Resolved group must have a material attached to it");
    }
    geometry.insert(geometry.end(), units.begin(), units.end());
    return true;
  }

  // ---------------------------------------------------------------- camera and renderer
  // parseCamera (:476-514); SimpleStochasticLense::fromDson (SimpleStochasticLense.impl.cuh:28-43)
  bool parse_camera(const DValue &d) {
    if (d.kind != DValue::DICT) return error("Error: Camera should be contained in Dson::Dict");
    if (d.has("lense")) {
      const DValue &l = d.at("lense");
      if (l.kind != DValue::DICT) return error("Error: Camera lense has to be a Dson::Dict");
      if (!l.has("type")) return error("Error: Lense has to have a type");
      const DValue &t = l.at("type");
      if (t.kind != DValue::STRING) return error("Error: Lense type has to be a string");
      if (t.str == "simple_stochastic") {
        float x = lens_focal(60.0f);
        if (l.has("angle")) {
          if (l.at("angle").kind != DValue::NUMBER) return error("Error: SimpleStochasticLense angle can only be a number...");
          x = lens_focal(l.at("angle").f());
        }
        lens_set = true;
        lens_x = x;
        settings.lens = DR_STOCHASTIC_LENS;
        std::memset(settings.lens_p, 0, sizeof(settings.lens_p));
      } else if (t.str == "spherical_segment") {
        // Lenses::SphericalSegmentLense defaults (SphericalSegmentLense.cuh:9) and fromDson (.impl.cuh:38-70)
        float ax = 60.0f, ay = 60.0f, dist = 128.0f, soft = 1.0f;
        F3 filter = f3(1.0f, 1.0f, 1.0f);
        if (l.has("angle")) {
          if (l.at("angle").kind != DValue::NUMBER) return error("SimpleFocalPlaneLense angle must be a number");
          ax = ay = l.at("angle").f();
        }
        if (!number_of(l, "distance", dist, "SimpleFocalPlaneLense distance must be a number")) return false;
        if (!number_of(l, "sensor", soft, "SimpleFocalPlaneLense sensor size must be a number")) return false;
        if (l.has("filter") && !vec_from(l.at("filter"), filter, err)) return false;
        lens_set = true;
        settings.lens = DR_SPHERICAL_SEGMENT_LENS;
        const float lp[8] = {ax, ay, dist, soft, filter.x, filter.y, filter.z, 0.0f};
        std::memcpy(settings.lens_p, lp, sizeof(lp));
      } else {
        return error("Error: Unknown lense type: \"" + t.str + "\"");
      }
    }
    if (d.has("transform") && !transform_from(d.at("transform"), camera, err)) return false;
    return true;
  }

  // parseRenderer (:515-640).  Thread, device and block settings are validated and
  // otherwise ignored: this library schedules its own work.
  bool parse_renderer(const DValue &d) {
    if (d.kind != DValue::DICT) return error("Error: Renderer should be contained in Dson::Dict");
    if (d.has("resources")) {
      const DValue &r = d.at("resources");
      if (r.kind != DValue::DICT) return error("Error: Renderer Resources has to be a Dson::Dict type");
      if (r.has("cpu")) {
        const DValue &c = r.at("cpu");
        const bool ok = (c.kind == DValue::NUMBER) ||
                        (c.kind == DValue::STRING && (c.str == "all" || c.str == "all_but_gpu" ||
                                                      c.str == "all_but_one_per_gpu" || c.str == "none" || c.str == "one"));
        if (!ok)
          return error("Error: Renderer Resources CPU can only be one of 'all'/['all_but_gpu']/'all_but_one_per_gpu'/'none'/'one'/any_number");
      }
      if (r.has("gpu") && r.at("gpu").kind != DValue::NUMBER) return error("Error: Renderer Resources GPU can only be a number");
    }
    if (d.has("blocks")) {
      const DValue &b = d.at("blocks");
      if (b.kind != DValue::DICT) return error("Error: Renderer Blocks has to be a Dson::Dict type");
      if (b.has("cpu_cut") && b.at("cpu_cut").kind != DValue::NUMBER) return error("Error: Renderer Blocks CPU_CUT has to be a number");
      if (b.has("gpu_cut") && b.at("gpu_cut").kind != DValue::NUMBER) return error("Error: Renderer Blocks GPU_CUT has to be a number");
      if (b.has("force_host_block_synchronisation") && b.at("force_host_block_synchronisation").kind != DValue::BOOL)
        return error("Error: Renderer Blocks force_host_block_synchronisation has to be a boolean");
    }
    if (d.has("pixel")) {
      const DValue &p = d.at("pixel");
      if (p.kind != DValue::DICT) return error("Error: Renderer Pixel has to be a Dson::Dict type");
      if (p.has("boxing")) {
        const DValue &b = p.at("boxing");
        if (b.kind != DValue::STRING) return error("Error: Renderer Pixel boxing type has to be a string");
        if (b.str == "height") settings.boxing = DR_BOXING_HEIGHT;
        else if (b.str == "width") settings.boxing = DR_BOXING_WIDTH;
        else if (b.str == "min") settings.boxing = DR_BOXING_MIN;
        else if (b.str == "max") settings.boxing = DR_BOXING_MAX;
        else return error("Error: Renderer Pixel boxing mode can only be one of ['height']/'width'/'min'/'max'");
      }
      if (p.has("bounces")) {
        if (p.at("bounces").kind != DValue::NUMBER) return error("Error: Renderer Pixel bounces has to be a number");
        settings.max_bounces = p.at("bounces").i();
      }
      if (p.has("samples_per_pixel")) {
        const DValue &s = p.at("samples_per_pixel");
        if (s.kind != DValue::ARRAY || s.arr.size() != 2 || s.arr[0]->kind != DValue::NUMBER || s.arr[1]->kind != DValue::NUMBER)
          return error("Error: Renderer Pixel samples_per_pixel has to be an array, containing two numbers");
        settings.fsaa_x = s.arr[0]->i();
        settings.fsaa_y = s.arr[1]->i();
      }
      if (p.has("pixels_per_gpu_thread") && p.at("pixels_per_gpu_thread").kind != DValue::NUMBER)
        return error("Error: Renderer Pixel pixels_per_gpu_thread has to be a number");
      if (p.has("ignore_backfaces")) {
        if (p.at("ignore_backfaces").kind != DValue::BOOL) return error("Error: Renderer Pixel ignore_backfaces has to have a boolean value");
        settings.ignore_backfaces = p.at("ignore_backfaces").b ? 1 : 0;
      }
    }
    return true;
  }

  void fill_camera(dr_scene_file_settings &st) const {
    const F3 o = camera.point(f3(0.0f, 0.0f, 0.0f));   // Camera::getPixelSamples: lens ray origin 0 moved by the transform
    const F3 r = camera.right(), u = camera.up(), f = camera.front();
    const float v[13] = {o.x, o.y, o.z, r.x, r.y, r.z, u.x, u.y, u.z, f.x, f.y, f.z, lens_x};
    std::memcpy(st.cam, v, sizeof(v));
    const float rows[9] = {camera.tx.x, camera.tx.y, camera.tx.z, camera.ty.x, camera.ty.y,
                           camera.ty.z, camera.tz.x, camera.tz.y, camera.tz.z};
    std::memcpy(st.cam_rows, rows, sizeof(rows));
  }
};

namespace {
int finish_load(std::unique_ptr<dr_scene_file_s> f, bool ok, dr_scene_file *out) {
  if (!ok) return dr::fail(DR_ERR_ARG, f->err.empty() ? std::string("scene file could not be loaded") : f->err);
  *out = f.release();
  return DR_OK;
}
}  // namespace

extern "C" {

int dr_scene_file_load(const char *path, dr_scene_file *out) {
  if (!path || !out) return dr::fail(DR_ERR_ARG, "dr_scene_file_load: bad arguments");
  *out = nullptr;
  try {
    auto f = std::make_unique<dr_scene_file_s>();   // freed here if loading throws
    const bool ok = f->from_file(path, 0);
    return finish_load(std::move(f), ok, out);
  } catch (const std::bad_alloc &) {
    return dr::fail(DR_ERR_ALLOC, "dr_scene_file_load: out of host memory");
  } catch (const std::exception &e) {
    return dr::fail(DR_ERR_ARG, std::string("dr_scene_file_load: ") + e.what());
  }
}

int dr_scene_file_parse(const char *text, const char *base_dir, dr_scene_file *out) {
  if (!text || !out) return dr::fail(DR_ERR_ARG, "dr_scene_file_parse: bad arguments");
  *out = nullptr;
  try {
    // names are appended to the directory as the reference appends them to its source
    // path, which ends in a separator
    std::string dir = base_dir ? base_dir : "";
    if (!dir.empty() && dir.back() != '/' && dir.back() != '\\') dir += '/';
    auto f = std::make_unique<dr_scene_file_s>();
    const bool ok = f->from_text(text, dir, 0);
    return finish_load(std::move(f), ok, out);
  } catch (const std::bad_alloc &) {
    return dr::fail(DR_ERR_ALLOC, "dr_scene_file_parse: out of host memory");
  } catch (const std::exception &e) {
    return dr::fail(DR_ERR_ARG, std::string("dr_scene_file_parse: ") + e.what());
  }
}

int dr_scene_file_info(dr_scene_file f, dr_scene_file_settings *settings, int *ntris, int *nmat, int *nlights) {
  if (!f) return dr::fail(DR_ERR_ARG, "dr_scene_file_info: null scene file");
  if (settings) {
    *settings = f->settings;
    f->fill_camera(*settings);
  }
  if (ntris) *ntris = (int)f->geometry.size();
  if (nmat) *nmat = (int)f->materials.size();
  if (nlights) *nlights = (int)f->lights.size();
  return DR_OK;
}

int dr_scene_file_geometry(dr_scene_file f, float *tris, float *normals, int32_t *mat_id) {
  if (!f) return dr::fail(DR_ERR_ARG, "dr_scene_file_geometry: null scene file");
  for (size_t i = 0; i < f->geometry.size(); i++) {
    const Unit &u = f->geometry[i];
    for (int k = 0; k < 3; k++) {
      if (tris) { tris[9 * i + 3 * k] = u.v[k].x; tris[9 * i + 3 * k + 1] = u.v[k].y; tris[9 * i + 3 * k + 2] = u.v[k].z; }
      if (normals) {
        normals[9 * i + 3 * k] = u.n[k].x; normals[9 * i + 3 * k + 1] = u.n[k].y; normals[9 * i + 3 * k + 2] = u.n[k].z;
      }
    }
    if (mat_id) mat_id[i] = u.mat;
  }
  return DR_OK;
}

int dr_scene_file_texture_info(dr_scene_file f, int *ntextures, int64_t *ntexels, int *nmaps) {
  if (!f) return dr::fail(DR_ERR_ARG, "dr_scene_file_texture_info: null scene file");
  if (ntextures) *ntextures = (int)f->textures.size();
  if (ntexels) *ntexels = (int64_t)(f->texels.size() / 4);
  if (nmaps) *nmaps = (int)f->maps.size();
  return DR_OK;
}

int dr_scene_file_textures(dr_scene_file f, dr_texture *textures, float *texels, dr_material_maps *maps) {
  if (!f) return dr::fail(DR_ERR_ARG, "dr_scene_file_textures: null scene file");
  if (textures && !f->textures.empty()) std::memcpy(textures, f->textures.data(), f->textures.size() * sizeof(dr_texture));
  if (texels && !f->texels.empty()) std::memcpy(texels, f->texels.data(), f->texels.size() * sizeof(float));
  if (maps && !f->maps.empty()) std::memcpy(maps, f->maps.data(), f->maps.size() * sizeof(dr_material_maps));
  return DR_OK;
}

int dr_scene_file_texcoords(dr_scene_file f, float *uv) {
  if (!f || !uv) return dr::fail(DR_ERR_ARG, "dr_scene_file_texcoords: bad arguments");
  for (size_t i = 0; i < f->geometry.size(); i++)
    for (int k = 0; k < 3; k++) {
      uv[6 * i + 2 * k] = f->geometry[i].t[k].x;
      uv[6 * i + 2 * k + 1] = f->geometry[i].t[k].y;
    }
  return DR_OK;
}

int dr_scene_file_tables(dr_scene_file f, dr_material *materials, dr_light *lights) {
  if (!f) return dr::fail(DR_ERR_ARG, "dr_scene_file_tables: null scene file");
  if (materials && !f->materials.empty()) std::memcpy(materials, f->materials.data(), f->materials.size() * sizeof(dr_material));
  if (lights && !f->lights.empty()) std::memcpy(lights, f->lights.data(), f->lights.size() * sizeof(dr_light));
  return DR_OK;
}

void dr_scene_file_destroy(dr_scene_file f) { delete f; }

}  // extern "C"
