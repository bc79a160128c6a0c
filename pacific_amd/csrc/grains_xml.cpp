// Grains3D XML input files -> a DEM system (grains_gpu_create_from_xml).
//
// A small self-contained XML reader (elements, attributes, text, comments,
// processing instructions, DOCTYPE, CDATA, the five predefined entities)
// replaces the Xerces DOM of GRAINS_V1/Grains/XML/src/ReaderXML.cpp; the
// schema mapping follows the code that consumes the DOM in the reference:
//   Grains::Construction ......... Base/src/Grains.cpp:535-1033
//   Grains::Forces ............... Base/src/Grains.cpp:1039-1100
//   Grains::AdditionalFeatures ... Base/src/Grains.cpp:1107-1960
//   ContactBuilderFactory::define  Component/src/ContactBuilderFactory.cpp:82-172
//   HookeContactForceModel::defineParameters  HookeContactForceModel.cpp:170-199
//   Transform(DOMNode*) .......... Geometry/src/Transform.cpp:490-535
// Supported: spheres, box obstacles (plain or inside <Composite>), Hooke
// models, insertion at the initial time in file order from a position file or
// a structured array, Zero / Constant / Random initial velocities.  Anything
// else is rejected with GRAINS_CONFIG_ERROR and a message naming the node.
#include <hip/hip_runtime.h>

#include <cctype>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/grains_gpu.h"
#include "grains_insert.hpp"
#include "grains_xml.hpp"
#include "dem_system.hpp"

namespace {

struct XNode {
  std::string name;
  std::vector<std::pair<std::string, std::string>> attrs;
  std::vector<XNode> kids;
  std::string text;

  const XNode* child(const char* n) const {
    for (auto& k : kids)
      if (k.name == n) return &k;
    return nullptr;
  }
  std::vector<const XNode*> children(const char* n) const {
    std::vector<const XNode*> out;
    for (auto& k : kids)
      if (k.name == n) out.push_back(&k);
    return out;
  }
  bool has(const char* a) const {
    for (auto& p : attrs)
      if (p.first == a) return true;
    return false;
  }
  std::string attr(const char* a) const {
    for (auto& p : attrs)
      if (p.first == a) return p.second;
    return std::string();
  }
};

struct XmlError {
  std::string msg;
};

class Parser {
 public:
  explicit Parser(const std::string& s) : s_(s), i_(0) {}

  XNode document() {
    XNode root;
    skip_misc();
    if (i_ >= s_.size() || s_[i_] != '<') fail("no root element");
    element(root);
    skip_misc();
    return root;
  }

 private:
  const std::string& s_;
  size_t i_;

  [[noreturn]] void fail(const std::string& m) {
    size_t line = 1;
    for (size_t k = 0; k < i_ && k < s_.size(); ++k) line += s_[k] == '\n';
    throw XmlError{"XML line " + std::to_string(line) + ": " + m};
  }
  bool starts(const char* p) const { return s_.compare(i_, std::strlen(p), p) == 0; }
  void skip_ws() {
    while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) ++i_;
  }
  void skip_until(const char* end) {
    size_t k = s_.find(end, i_);
    if (k == std::string::npos) fail(std::string("unterminated construct, expected ") + end);
    i_ = k + std::strlen(end);
  }
  void skip_doctype() {
    i_ += 9;  // "<!DOCTYPE"
    int depth = 0;
    while (i_ < s_.size()) {
      char c = s_[i_++];
      if (c == '[') ++depth;
      else if (c == ']') --depth;
      else if (c == '>' && depth <= 0) return;
    }
    fail("unterminated DOCTYPE");
  }
  // whitespace, comments, processing instructions, DOCTYPE
  void skip_misc() {
    for (;;) {
      skip_ws();
      if (starts("<?")) skip_until("?>");
      else if (starts("<!--")) skip_until("-->");
      else if (starts("<!DOCTYPE")) skip_doctype();
      else return;
    }
  }
  std::string name() {
    size_t b = i_;
    while (i_ < s_.size() && (std::isalnum((unsigned char)s_[i_]) || s_[i_] == '_' || s_[i_] == '-' ||
                              s_[i_] == '.' || s_[i_] == ':'))
      ++i_;
    if (b == i_) fail("expected a name");
    return s_.substr(b, i_ - b);
  }
  static std::string decode(const std::string& t) {
    std::string o;
    for (size_t k = 0; k < t.size(); ++k) {
      if (t[k] != '&') { o += t[k]; continue; }
      size_t e = t.find(';', k);
      if (e == std::string::npos) { o += t[k]; continue; }
      std::string ent = t.substr(k + 1, e - k - 1);
      if (ent == "lt") o += '<';
      else if (ent == "gt") o += '>';
      else if (ent == "amp") o += '&';
      else if (ent == "quot") o += '"';
      else if (ent == "apos") o += '\'';
      else if (!ent.empty() && ent[0] == '#') {
        long c = ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X') ? std::strtol(ent.c_str() + 2, nullptr, 16)
                                                                      : std::strtol(ent.c_str() + 1, nullptr, 10);
        if (c > 0 && c < 128) o += (char)c;
      } else {
        o += t.substr(k, e - k + 1);
      }
      k = e;
    }
    return o;
  }
  void element(XNode& n) {
    ++i_;  // '<'
    n.name = name();
    for (;;) {
      skip_ws();
      if (i_ >= s_.size()) fail("unterminated tag <" + n.name + ">");
      if (starts("/>")) { i_ += 2; return; }
      if (s_[i_] == '>') { ++i_; break; }
      std::string a = name();
      skip_ws();
      if (i_ >= s_.size() || s_[i_] != '=') fail("expected '=' after attribute " + a);
      ++i_;
      skip_ws();
      if (i_ >= s_.size() || (s_[i_] != '"' && s_[i_] != '\'')) fail("expected a quoted value for " + a);
      char q = s_[i_++];
      size_t e = s_.find(q, i_);
      if (e == std::string::npos) fail("unterminated attribute value");
      n.attrs.push_back({a, decode(s_.substr(i_, e - i_))});
      i_ = e + 1;
    }
    // content
    for (;;) {
      if (i_ >= s_.size()) fail("element <" + n.name + "> not closed");
      if (starts("</")) {
        i_ += 2;
        std::string nm = name();
        if (nm != n.name) fail("closing </" + nm + "> does not match <" + n.name + ">");
        skip_ws();
        if (i_ >= s_.size() || s_[i_] != '>') fail("expected '>'");
        ++i_;
        return;
      }
      if (starts("<!--")) { skip_until("-->"); continue; }
      if (starts("<?")) { skip_until("?>"); continue; }
      if (starts("<![CDATA[")) {
        i_ += 9;
        size_t e = s_.find("]]>", i_);
        if (e == std::string::npos) fail("unterminated CDATA");
        n.text += s_.substr(i_, e - i_);
        i_ = e + 3;
        continue;
      }
      if (s_[i_] == '<') {
        n.kids.emplace_back();
        element(n.kids.back());
        continue;
      }
      size_t e = s_.find('<', i_);
      if (e == std::string::npos) e = s_.size();
      n.text += decode(s_.substr(i_, e - i_));
      i_ = e;
    }
  }
};

std::string trim(const std::string& s) {
  size_t b = 0, e = s.size();
  while (b < e && std::isspace((unsigned char)s[b])) ++b;
  while (e > b && std::isspace((unsigned char)s[e - 1])) --e;
  return s.substr(b, e - b);
}

struct Cfg {
  std::string err;
  bool ok = true;
  void bad(const std::string& m) {
    if (ok) err = m;
    ok = false;
  }
};

// ReaderXML::getNodeAttr_Double / getNodeValue_Double
double dattr(const XNode* n, const char* a, Cfg& c) {
  if (!n || !n->has(a)) {
    c.bad(std::string("missing attribute ") + a + (n ? " in <" + n->name + ">" : std::string()));
    return 0.;
  }
  std::string v = trim(n->attr(a));
  char* end = nullptr;
  double d = std::strtod(v.c_str(), &end);
  if (end == v.c_str()) c.bad("attribute " + std::string(a) + " of <" + n->name + "> is not a number");
  return d;
}
double dvalue(const XNode* n, Cfg& c, const char* what) {
  if (!n) {
    c.bad(std::string("missing node <") + what + ">");
    return 0.;
  }
  std::string v = trim(n->text);
  char* end = nullptr;
  double d = std::strtod(v.c_str(), &end);
  if (end == v.c_str()) c.bad("<" + n->name + "> does not hold a number");
  return d;
}

// <AngularPosition Type="Identity|Matrix">  (Transform.cpp:510-530)
bool basis_of(const XNode* owner, double B[9], Cfg& c) {
  for (int i = 0; i < 9; ++i) B[i] = (i % 4) == 0 ? 1. : 0.;
  const XNode* ap = owner->child("AngularPosition");
  if (!ap) return true;
  std::string mode = ap->attr("Type");
  if (mode == "Matrix") {
    std::istringstream in(ap->text);
    for (int i = 0; i < 9; ++i)
      if (!(in >> B[i])) {
        c.bad("AngularPosition matrix needs 9 values");
        return false;
      }
    // Matrix::isRotation: orthonormal with determinant +1
    double det = B[0] * (B[4] * B[8] - B[5] * B[7]) - B[1] * (B[3] * B[8] - B[5] * B[6]) +
                 B[2] * (B[3] * B[7] - B[4] * B[6]);
    if (std::fabs(det - 1.) > 1e-8) {
      c.bad("a matrix in one of the AngularPosition nodes is not a rotation matrix");
      return false;
    }
  }
  return true;
}

struct Material {
  std::map<std::string, int> index;
  int define(const std::string& name, Cfg& c) {  // ContactBuilderFactory::defineMaterial
    auto it = index.find(name);
    if (it != index.end()) return it->second;
    int k = (int)index.size();
    if (k >= dem::kMaxMaterials) c.bad("too many materials");
    index[name] = k;
    return k;
  }
};

// One <Particle> class: a sphere, a box or a polyhedron with its crust and
// the reference particle's angular position (inherited by inserted
// particles, InitialAngularPosition Type="Fixed")
struct ParticleClass {
  std::string file;  // polyhedron file name as given in the input
  int number;
  double density, crust, radius;
  int type;  // dem::ShapeKind
  double L[3];
  std::vector<double> verts;
  std::vector<int> face_sizes, face_idx;
  double B[9];
  int mat;
};

// Polyhedron file (Polyhedron::create, Polyhedron.cpp:100-160, readFaces
// :363-405): dimension, vertex count, vertices, face count, one face per line
bool read_polyhedron(const std::string& path, ParticleClass& pc, std::string& err) {
  std::ifstream in(path);
  if (!in) {
    err = "cannot open polyhedron file " + path;
    return false;
  }
  int dim = 0, nv = 0;
  in >> dim >> nv;
  if (!in || nv < 4) {
    err = "bad polyhedron header in " + path;
    return false;
  }
  pc.verts.resize(3 * (size_t)nv);
  for (int i = 0; i < 3 * nv; ++i)
    if (!(in >> pc.verts[i])) {
      err = "bad polyhedron vertices in " + path;
      return false;
    }
  int nf = 0;
  in >> nf;
  std::string line;
  std::getline(in, line);
  for (int f = 0; f < nf && std::getline(in, line);) {
    std::istringstream ls(line);
    int v, m = 0;
    while (ls >> v) {
      pc.face_idx.push_back(v);
      ++m;
    }
    if (m == 0) continue;  // blank line
    pc.face_sizes.push_back(m);
    ++f;
  }
  if ((int)pc.face_sizes.size() != nf) {
    err = "bad polyhedron faces in " + path;
    return false;
  }
  return true;
}

}  // namespace

extern "C" int grains_gpu_create_from_xml(const char* xml_path, int precision, grains_gpu_system** out) {
  return dem::create_from_xml(xml_path, precision, dem::XmlOptions{}, out);
}

extern "C" int grains_gpu_parse_xml(const char* xml_path, int precision, grains_gpu_system** out) {
  dem::XmlOptions opt;
  opt.no_init = true;
  return dem::create_from_xml(xml_path, precision, opt, out);
}

int dem::create_from_xml(const char* xml_path, int precision, const XmlOptions& opt, grains_gpu_system** out) {
  if (!xml_path || !out) return GRAINS_ARGUMENT_ERROR;
  *out = nullptr;
  std::ifstream f(xml_path);
  if (!f) {
    grains_gpu_config cfg0{};
    grains_gpu_create(&cfg0, out);
    (*out)->last_error = std::string("cannot open ") + xml_path;
    return GRAINS_CONFIG_ERROR;
  }
  std::stringstream buf;
  buf << f.rdbuf();
  std::string src = buf.str();

  grains_gpu_config cfg{};
  cfg.cell_size_factor = 1.;
  cfg.precision = precision;
  cfg.scheme = GRAINS_LEAPFROG;
  Cfg c;
  XNode root;
  try {
    root = Parser(src).document();
  } catch (XmlError& e) {
    c.bad(e.msg);
  }
  std::string dir;
  {
    std::string p(xml_path);
    size_t k = p.find_last_of('/');
    dir = k == std::string::npos ? std::string() : p.substr(0, k + 1);
  }

  Material mats;
  std::vector<ParticleClass> classes;
  struct BoxDef {
    double L[3], centre[3], B[9], crust;
    int mat, to_fluid;
    std::string name;
  };
  std::vector<BoxDef> boxes;
  struct ModelDef {
    int a, b;
    double kn, en, etat, muc, kr;
    bool memory = false;  // <HookeMemory>: kt, mur, etarpf, epst
    double kt = 0., mur = 0., etarpf = 0., epst = 0.;
  };
  std::vector<ModelDef> models;
  std::vector<double> positions;
  grains_insert::Plan plan;
  unsigned seed = 0;
  double t_end = 0.;
  std::vector<double> save_times;
  std::string save_root, rawdata_root;
  bool rawdata_binary = false;
  double coupled_dt[2] = {0., 0.};
  size_t coupled_ndt = 0;
  std::string reload_type, reload_file, reload_velocity;

  if (c.ok && root.name != "Grains3D") c.bad("root element must be <Grains3D>");
  if (c.ok && !opt.coupled && root.has("Type") && root.attr("Type") != "Standard")
    c.bad("only <Grains3D Type=\"Standard\"> is supported");
  cfg.fluid_density = opt.fluid_density;
  const XNode* cons = c.ok ? root.child("Construction") : nullptr;
  dem::GjkOpts gjk{0, 0, 0, 0, 1e-10};
  if (c.ok && !cons) c.bad("missing <Construction>");

  // ---- Construction (Grains.cpp:535-1033)
  if (c.ok) {
    const XNode* lc = cons->child("LinkedCell");
    cfg.max[0] = dattr(lc, "MX", c);
    cfg.max[1] = dattr(lc, "MY", c);
    cfg.max[2] = dattr(lc, "MZ", c);
    if (lc && lc->has("CellSizeFactor")) cfg.cell_size_factor = dattr(lc, "CellSizeFactor", c);
    if (const XNode* o = cons->child("Origin")) {
      cfg.origin[0] = dattr(o, "OX", c);
      cfg.origin[1] = dattr(o, "OY", c);
      cfg.origin[2] = dattr(o, "OZ", c);
    }
    if (const XNode* p = cons->child("Periodicity")) {
      const char* ax[3] = {"PX", "PY", "PZ"};
      for (int a = 0; a < 3; ++a) cfg.periodic[a] = std::atoi(p->attr(ax[a]).c_str()) == 1 ? 1 : 0;
    }
    if (cons->child("PartialPeriodicity")) c.bad("<PartialPeriodicity> is not supported");
    // <CollisionDetectionAlgorithm> (Grains.cpp:640-770): <CollisionDetection
    // Type="GJK|GJK_SV" Tolerance Acceleration="ON|OFF" History="OFF"> and
    // <BoundingVolume Type="OBB|OBC">.  History="ON" (the warm-started GJK of
    // Convex.cpp:636-733) is refused rather than ignored.
    if (const XNode* cda = cons->child("CollisionDetectionAlgorithm")) {
      if (const XNode* cd = cda->child("CollisionDetection")) {
        if (cd->has("Type")) {
          const std::string t = cd->attr("Type");
          if (t == "GJK_SV") gjk.sv = 1;
          else if (t != "GJK") c.bad("collision detection algorithm " + t + " is not defined");
        }
        if (c.ok && cd->has("Tolerance")) {
          const double tol = dattr(cd, "Tolerance", c);
          if (c.ok && tol < 1e-10) c.bad("GJK tolerance should be at least 1E-10");
          gjk.tol = tol;
        }
        if (c.ok && cd->has("Acceleration")) {
          const std::string v = cd->attr("Acceleration");
          if (v == "ON") gjk.accel = 1;
          else if (v != "OFF") c.bad("Acceleration should be ON or OFF");
        }
        if (c.ok && cd->has("History")) {
          const std::string v = cd->attr("History");
          if (v == "ON") c.bad("GJK History=\"ON\" is not supported");
          else if (v != "OFF") c.bad("History should be ON or OFF");
        }
      }
      if (const XNode* bvn = cda->child("BoundingVolume")) {
        const std::string t = bvn->attr("Type");
        gjk.bv = t == "OBB" ? 1 : (t == "OBC" ? 2 : 0);
      }
    }
    // <Reload Type="new" Filename="..."> or <Reload Type="same"> (the last
    // file of <RestartFile Name>_RFTable.txt), Velocity="Reset" optional
    // (Grains.cpp:781-844): the particles the input file inserts take the
    // state of the reload file (same ids and classes); obstacles and
    // contact models are those of the input file
    if (const XNode* rl = cons->child("Reload")) {
      reload_type = rl->attr("Type");
      if (reload_type != "new" && reload_type != "same") c.bad("<Reload Type> must be \"new\" or \"same\"");
      reload_file = rl->attr("Filename");
      reload_velocity = rl->attr("Velocity");
      if (c.ok && reload_type == "new" && reload_file.empty()) c.bad("<Reload Type=\"new\"> needs a Filename");
    }
    if (cons->child("CompositeParticles")) c.bad("<CompositeParticles> are not supported");
    if (const XNode* ps = cons->child("Particles")) {
      for (const XNode* p : ps->children("Particle")) {
        ParticleClass sc;
        sc.number = std::atoi(p->attr("Number").c_str());
        sc.density = dattr(p, "Density", c);
        sc.radius = 0.;
        sc.L[0] = sc.L[1] = sc.L[2] = 0.;
        const XNode* cv = p->child("Convex");
        if (!cv) { c.bad("<Particle> without <Convex>"); break; }
        sc.crust = dattr(cv, "CrustThickness", c);
        if (const XNode* sph = cv->child("Sphere")) {
          sc.type = dem::kShapeSphere;
          sc.radius = dattr(sph, "Radius", c);
        } else if (const XNode* bx = cv->child("Box")) {
          sc.type = dem::kShapeBox;
          sc.L[0] = dattr(bx, "LX", c);
          sc.L[1] = dattr(bx, "LY", c);
          sc.L[2] = dattr(bx, "LZ", c);
        } else if (const XNode* ph = cv->child("Polyhedron")) {
          sc.type = dem::kShapePoly;
          std::string name = ph->attr("Name"), err;
          sc.file = name;
          std::ifstream probe(name);
          if (!probe) name = dir + name;
          if (!read_polyhedron(name, sc, err)) { c.bad(err); break; }
        } else {
          c.bad("particle convexes other than <Sphere>, <Box> and <Polyhedron> are not supported");
          break;
        }
        basis_of(p, sc.B, c);
        const XNode* m = p->child("Material");
        sc.mat = mats.define(m ? trim(m->text) : std::string(""), c);
        classes.push_back(sc);
      }
    }
    if (const XNode* os = cons->child("Obstacles")) {
      // ObstacleBuilderFactory::create: <Composite> nests obstacles
      std::vector<const XNode*> stack;
      for (auto it = os->kids.rbegin(); it != os->kids.rend(); ++it) stack.push_back(&*it);
      std::vector<const XNode*> flat;
      while (!stack.empty()) {
        const XNode* n = stack.back();
        stack.pop_back();
        if (n->name == "Composite") {
          for (auto it = n->kids.rbegin(); it != n->kids.rend(); ++it)
            if (it->name == "Composite" || it->name == "Obstacle") stack.push_back(&*it);
        } else if (n->name == "Obstacle") {
          flat.push_back(n);
        } else {
          c.bad("obstacle type <" + n->name + "> is not supported");
        }
      }
      for (const XNode* n : flat) {
        BoxDef b;
        const XNode* cv = n->child("Convex");
        if (!cv) { c.bad("<Obstacle> without <Convex>"); break; }
        b.crust = dattr(cv, "CrustThickness", c);
        const XNode* bx = cv->child("Box");
        if (!bx) { c.bad("only <Box> obstacles are supported"); break; }
        b.L[0] = dattr(bx, "LX", c);
        b.L[1] = dattr(bx, "LY", c);
        b.L[2] = dattr(bx, "LZ", c);
        const XNode* ce = n->child("Centre");
        b.centre[0] = dattr(ce, "X", c);
        b.centre[1] = dattr(ce, "Y", c);
        b.centre[2] = dattr(ce, "Z", c);
        basis_of(n, b.B, c);
        const XNode* m = n->child("Material");
        b.mat = mats.define(m ? trim(m->text) : std::string(""), c);
        b.name = n->attr("name");
        // <Status ToFluid="1"/> (SimpleObstacle.cpp:54-57)
        const XNode* stn = n->child("Status");
        b.to_fluid = stn ? std::atoi(stn->attr("ToFluid").c_str()) : 0;
        boxes.push_back(b);
      }
    }
    if (const XNode* cm = cons->child("ContactForceModels")) {
      for (const XNode& k : cm->kids) {
        ModelDef md;
        const XNode* law = nullptr;
        if (k.name == "Default") {
          md.a = md.b = -1;
        } else if (k.name == "ContactForceModel") {
          const XNode* m = k.child("Material");
          if (!m) { c.bad("<ContactForceModel> without <Material>"); break; }
          md.a = mats.define(m->attr("materialA"), c);
          md.b = mats.define(m->attr("materialB"), c);
        } else {
          continue;
        }
        for (const XNode& t : k.kids) {
          if (t.name == "Material") continue;
          if (t.name == "Hooke" || t.name == "HookeMemory") law = &t;
          else c.bad("contact force model <" + t.name + "> is not supported (only <Hooke> and <HookeMemory>)");
        }
        if (!law) { c.bad("no contact force model in <" + k.name + ">"); break; }
        md.kn = dvalue(law->child("kn"), c, "kn");
        md.en = dvalue(law->child("en"), c, "en");
        md.muc = dvalue(law->child("muc"), c, "muc");
        md.etat = law->child("etat") ? dvalue(law->child("etat"), c, "etat") : -1.;
        if (law->name == "HookeMemory") {
          // HookeMemoryContactForceModel::defineParameters (:253-307); a
          // missing <epst> leaves the threshold at 0 there (the default is
          // stored under another key)
          md.memory = true;
          md.kr = 0.;
          md.kt = dvalue(law->child("kt"), c, "kt");
          md.mur = law->child("mur") ? dvalue(law->child("mur"), c, "mur") : 0.;
          md.etarpf = law->child("etarpf") ? dvalue(law->child("etarpf"), c, "etarpf") : 0.;
          md.epst = law->child("epst") ? dvalue(law->child("epst"), c, "epst") : 0.;
        } else {
          md.kr = dvalue(law->child("kr"), c, "kr");
        }
        models.push_back(md);
      }
    }
  }

  // ---- Forces (Grains.cpp:1039-1100): gravity is mandatory
  if (c.ok) {
    const XNode* fo = root.child("Forces");
    const XNode* g = fo ? fo->child("Gravity") : nullptr;
    if (!g) c.bad("missing <Forces><Gravity>");
    else {
      cfg.gravity[0] = dattr(g, "GX", c);
      cfg.gravity[1] = dattr(g, "GY", c);
      cfg.gravity[2] = dattr(g, "GZ", c);
    }
  }

  // ---- Simulation (Grains.cpp:1107-1960)
  if (c.ok) {
    const XNode* sim = root.child("Simulation");
    if (!sim) c.bad("missing <Simulation>");
    else {
      const XNode* ti = sim->child("TimeInterval");
      if (opt.coupled) {
        // GrainsCoupledWithFluid::Construction (GrainsCoupledWithFluid.cpp:
        // 509-527): <TimeStep mindt maxdt ndt>, no time interval
        const XNode* ts = sim->child("TimeStep");
        coupled_dt[0] = dattr(ts, "mindt", c);
        coupled_dt[1] = dattr(ts, "maxdt", c);
        coupled_ndt = (size_t)std::atol(ts ? ts->attr("ndt").c_str() : "0");
        if (c.ok && (!(coupled_dt[0] > 0.) || coupled_dt[1] < coupled_dt[0] || coupled_ndt == 0))
          c.bad("<TimeStep> needs 0 < mindt <= maxdt and ndt >= 1");
        cfg.dt = coupled_dt[0];
      } else {
        cfg.t_start = dattr(ti, "Start", c);
        t_end = dattr(ti, "End", c);
        cfg.dt = dattr(sim->child("TimeStep"), "dt", c);
      }
      // outputs: <TimeSave Start End Every> (Grains.cpp:1224-1235), the
      // <RestartFile Name> root (:1174-1182), RawData writers
      // (PostProcessingWriterBuilderFactory.cpp:14-20, RawDataPostProcessingWriter.cpp:20-30)
      if (const XNode* tsv = sim->child("TimeSave")) {
        const double t0 = dattr(tsv, "Start", c), t1 = dattr(tsv, "End", c);
        double every = dattr(tsv, "Every", c);
        if (every < cfg.dt) every = cfg.dt;
        if (c.ok && every > 0.)
          for (double t = t0; t - t1 < 0.01 * cfg.dt; t += every) save_times.push_back(t);
      }
      if (const XNode* rf = sim->child("RestartFile")) save_root = rf->attr("Name");
      if (const XNode* pp = sim->child("PostProcessing"))
        if (const XNode* wr = pp->child("Writers"))
          if (const XNode* rd = wr->child("RawData")) {
            rawdata_root = rd->attr("Name");
            rawdata_binary = rd->attr("WritingMode") == "Binary";
            if (rawdata_root.empty()) c.bad("<RawData> needs a Name");
          }
      if (const XNode* tint = sim->child("TimeIntegration")) {
        std::string t = tint->attr("Type");
        if (t == "SecondOrderLeapFrog") cfg.scheme = GRAINS_LEAPFROG;
        else if (t == "FirstOrderExplicit") cfg.scheme = GRAINS_FIRST_ORDER_EXPLICIT;
        else if (t == "SecondOrderExplicit") cfg.scheme = GRAINS_SECOND_ORDER_EXPLICIT;
        else if (t == "SecondOrderAdamsBashforth") cfg.scheme = GRAINS_ADAMS_BASHFORTH2;
        else c.bad("unknown time integrator " + t);
      }
      const XNode* ins = sim->child("ParticleInsertion");
      size_t nnew = 0;
      for (auto& sc : classes) nnew += (size_t)sc.number;
      if (nnew && !ins) c.bad("particles declared without <ParticleInsertion>");
      if (ins && c.ok) {
        // Grains.cpp:1281-1400
        if (const XNode* mode = ins->child("Mode")) {
          std::string t = mode->attr("Type");
          if (t == "OverTime") plan.over_time = true;
          else if (t != "InitialTime") c.bad("<Mode Type> is InitialTime or OverTime");
        }
        if (const XNode* order = ins->child("Order")) plan.random_order = order->attr("Type") == "Random";
        if (const XNode* iap = ins->child("InitialAngularPosition")) {
          plan.random_angle = iap->attr("Type") == "Random";
          if (iap->attr("Type") == "File") {
            // Grains::setAngularPositionParticlesFromFile (Grains.cpp:2256-2300):
            // one matrix by rows per particle to insert
            std::string name = iap->attr("Name");
            std::ifstream af(name);
            if (!af) af.open(dir + name);
            if (!af) c.bad("cannot open angular position file " + name);
            double m;
            while (af >> m) plan.angles.push_back(m);
            if (c.ok && plan.angles.size() != 9 * nnew)
              c.bad("the number of angular positions in " + name + " differs from the number of particles");
          }
        }
        if (const XNode* rgs = ins->child("RandomGeneratorSeed")) {
          std::string t = rgs->attr("Type");
          if (t == "UserDefined") seed = (unsigned)std::atoi(rgs->attr("Value").c_str());
          else if (t == "Random") c.bad("RandomGeneratorSeed Type=\"Random\" is not reproducible");
        }
        if (const XNode* fr = ins->child("Frequency")) {
          long k = std::atol(fr->attr("TryEvery").c_str());
          if (k < 1) c.bad("<Frequency TryEvery> must be at least 1");
          plan.frequency = (size_t)std::max(k, 1L);
        }
        if (const XNode* fi = ins->child("ForceInsertion")) plan.force = fi->attr("Value") == "True";
        if (const XNode* go = ins->child("GeometricOverlap")) plan.bv_only = go->attr("Value") == "BV";
        const XNode* pp = ins->child("ParticlePosition");
        if (!pp) c.bad("missing <ParticlePosition>");
        else if (const XNode* fl = pp->child("File")) {
          std::string name = fl->attr("Name");
          std::ifstream pf(name);
          if (!pf) pf.open(dir + name);
          if (!pf) c.bad("cannot open position file " + name);
          double x, y, z;
          while (pf >> x >> y >> z) {  // Point3 operator>> (Grains.cpp:2243)
            positions.push_back(x);
            positions.push_back(y);
            positions.push_back(z);
          }
        } else if (const XNode* st = pp->child("StructuredArray")) {
          // Grains.cpp:1421-1463 and setPositionParticlesStructuredArray :2303-2345
          const XNode* bx = st->child("Box");
          const XNode* nb = st->child("Number");
          if (!bx || bx->kids.size() < 2 || !nb) c.bad("<StructuredArray> needs <Box> with two points and <Number>");
          else {
            double A[3] = {dattr(&bx->kids[0], "X", c), dattr(&bx->kids[0], "Y", c), dattr(&bx->kids[0], "Z", c)};
            double B[3] = {dattr(&bx->kids[1], "X", c), dattr(&bx->kids[1], "Y", c), dattr(&bx->kids[1], "Z", c)};
            long NX = std::atol(nb->attr("NX").c_str()), NY = std::atol(nb->attr("NY").c_str()),
                 NZ = std::atol(nb->attr("NZ").c_str());
            if ((size_t)(NX * NY * NZ) != nnew) c.bad("structured array size differs from the number of particles");
            double dx = (B[0] - A[0]) / double(NX), dy = (B[1] - A[1]) / double(NY), dz = (B[2] - A[2]) / double(NZ);
            for (long k = 0; k < NX; ++k)
              for (long l = 0; l < NY; ++l)
                for (long m = 0; m < NZ; ++m) {
                  positions.push_back(A[0] + (double(k) + 0.5) * dx);
                  positions.push_back(A[1] + (double(l) + 0.5) * dy);
                  positions.push_back(A[2] + (double(m) + 0.5) * dz);
                }
          }
        } else if (const XNode* wins = pp->child("Windows")) {
          // Window::readWindow (Window.cpp:53-230)
          for (const XNode& wn : wins->kids) {
            grains_insert::Window w;
            std::string t = wn.has("Type") ? wn.attr("Type") : std::string("Box");
            auto point = [&](const XNode* n, double* P) {
              P[0] = dattr(n, "X", c);
              P[1] = dattr(n, "Y", c);
              P[2] = dattr(n, "Z", c);
            };
            auto axis = [&](const std::string& d) {
              if (d == "X") return 0;
              if (d == "Y") return 1;
              if (d != "Z") c.bad("window axis Direction is X, Y or Z");
              return 2;
            };
            if (t == "Box" || t == "Line") {
              w.type = t == "Box" ? grains_insert::kBox : grains_insert::kLine;
              if (wn.kids.size() < 2) {
                c.bad("a <Window Type=\"" + t + "\"> needs two points");
                break;
              }
              point(&wn.kids[0], w.A);
              point(&wn.kids[1], w.B);
            } else if (t == "Cylinder" || t == "Annulus") {
              const XNode* bc = wn.child("BottomCentre");
              const XNode* g = wn.child(t.c_str());
              if (!bc || !g) {
                c.bad("a <Window Type=\"" + t + "\"> needs <BottomCentre> and <" + t + ">");
                break;
              }
              point(bc, w.A);
              if (t == "Cylinder") {
                w.type = grains_insert::kCylinder;
                w.radius = dattr(g, "Radius", c);
              } else {
                w.type = grains_insert::kAnnulus;
                w.radius = dattr(g, "RadiusExt", c);
                w.radius_int = dattr(g, "RadiusInt", c);
              }
              w.height = dattr(g, "Height", c);
              w.axis = axis(g->attr("Direction"));
            } else {
              c.bad("unknown insertion window type " + t);
              break;
            }
            plan.windows.push_back(w);
          }
          if (plan.windows.empty()) c.bad("<Windows> holds no window");
        } else {
          c.bad("<ParticlePosition> needs <File>, <StructuredArray> or <Windows>");
        }
        if (c.ok && plan.windows.empty() && positions.size() / 3 < nnew)
          c.bad("fewer positions than particles to insert");
        if (const XNode* iv = ins->child("InitialVelocity")) {
          std::string m = iv->attr("Mode");
          if (m == "Constant") {
            plan.vmode = 1;
            if (const XNode* t = iv->child("TranslationalVelocity")) {
              plan.vconst[0] = dattr(t, "VX", c);
              plan.vconst[1] = dattr(t, "VY", c);
              plan.vconst[2] = dattr(t, "VZ", c);
            }
            if (const XNode* r = iv->child("AngularVelocity")) {
              plan.wconst[0] = dattr(r, "RX", c);
              plan.wconst[1] = dattr(r, "RY", c);
              plan.wconst[2] = dattr(r, "RZ", c);
            }
          } else if (m == "Random") {
            plan.vmode = 2;
            if (const XNode* t = iv->child("Translational")) plan.amp_t = dattr(t, "Amplitude", c);
            if (const XNode* r = iv->child("Angular")) plan.amp_r = dattr(r, "Amplitude", c);
          }
        }
      }
    }
  }

  grains_gpu_system* s = nullptr;
  grains_gpu_create(&cfg, &s);
  *out = s;
  if (!c.ok) {
    s->last_error = c.err;
    return GRAINS_CONFIG_ERROR;
  }
  s->gjk_opts = gjk;
  s->t_end = t_end;
  s->save_times = save_times;
  s->save_root = save_root;
  s->rawdata_root = rawdata_root;
  s->rawdata_binary = rawdata_binary;
  s->track_basis = opt.track_basis;
  s->fluid_corrected = opt.fluid_corrected;
  s->coupled_min_dt = coupled_dt[0];
  s->coupled_max_dt = coupled_dt[1];
  s->coupled_ndt = coupled_ndt;
  int e;
  for (auto& m : models) {
    if (m.memory) {
      const double p[8] = {m.kn, m.en, m.etat, m.muc, m.kt, m.mur, m.etarpf, m.epst};
      e = grains_gpu_set_memory_contact_model(s, m.a, m.b, p);
    } else {
      e = grains_gpu_set_contact_model(s, m.a, m.b, m.kn, m.en, m.etat, m.muc, m.kr);
    }
    if (e) return e;
  }
  // insertion at the initial time, class by class in file order
  // (AllComponents::getParticleToInsert PM_ORDERED, Grains.cpp:2070-2140)
  // shape classes for the non-spherical particles (spheres of a system
  // without them keep the sphere-only kernels)
  bool any_shape = false;
  for (auto& sc : classes) any_shape |= sc.type != dem::kShapeSphere;
  std::vector<int> shape_of(classes.size(), -1);
  for (size_t i = 0; i < classes.size(); ++i) {
    ParticleClass& sc = classes[i];
    bool identity = true;
    for (int j = 0; j < 9; ++j) identity &= sc.B[j] == ((j % 4) == 0 ? 1. : 0.);
    if (sc.type == dem::kShapeSphere && !any_shape && identity) continue;
    e = grains_gpu_add_shape(s, sc.type, sc.radius, sc.L, (int)(sc.verts.size() / 3), sc.verts.data(),
                             (int)sc.face_sizes.size(), sc.face_sizes.data(), sc.face_idx.data(), sc.crust,
                             &shape_of[i]);
    if (e) return e;
  }
  for (auto& b : boxes) {
    e = grains_gpu_add_box(s, b.L, b.centre, b.B, b.crust, b.mat);
    if (e) return e;
    s->obs_to_fluid.back() = b.to_fluid;
    s->obs_names.push_back(b.name);
  }
  // the particles of the input, inserted at the initial time or over time
  // (Grains.cpp:2004-2017, 322-327), class by class
  srand(seed);
  for (size_t i = 0; i < classes.size(); ++i) {
    const ParticleClass& sc = classes[i];
    grains_insert::Class k;
    k.shape = shape_of[i];
    k.xml_class = (int)i;
    k.radius = sc.radius;
    k.density = sc.density;
    k.crust = sc.crust;
    k.mat = sc.mat;
    for (int j = 0; j < 9; ++j) k.B[j] = sc.B[j];
    k.remaining = sc.number;
    plan.classes.push_back(k);
  }
  // ordered insertion reverses the list and takes the last entry
  // (Grains.cpp:1991-1995)
  for (size_t i = 0; !plan.random_order && i < positions.size() / 2 / 3; ++i)
    for (int a = 0; a < 3; ++a) std::swap(positions[3 * i + a], positions[positions.size() - 3 * (i + 1) + a]);
  plan.positions = positions;
  auto ins_plan = std::make_shared<grains_insert::Plan>(plan);
  if (!plan.over_time) {
    if ((e = grains_insert::insert_all(s, *ins_plan, 1000000))) return e;
  } else {
    s->insertion = ins_plan;
    // an initialisation needs a particle: the first steps' attempts are made
    // here while the system is empty (each one that fails is a step of an
    // empty system)
    while (s->nparticles == 0 && ins_plan->pending() > 0) {
      bool in = false;
      if ((e = grains_insert::attempt(s, *ins_plan, &in))) return e;
      if (in) ins_plan->preattempted = true;
      else s->time += cfg.dt;
    }
  }
  s->mat_names.assign(dem::kMaxMaterials, std::string());
  for (auto& kv : mats.index) s->mat_names[(size_t)kv.second] = kv.first;
  for (size_t i = 0; i < classes.size(); ++i)
    if (shape_of[i] >= 0) {
      if (s->shape_files.size() <= (size_t)shape_of[i]) s->shape_files.resize((size_t)shape_of[i] + 1);
      s->shape_files[(size_t)shape_of[i]] = classes[i].file;
    }
  if (opt.no_init) return GRAINS_OK;
  if ((e = grains_gpu_initialize(s))) return e;
  if (reload_type.empty()) return GRAINS_OK;
  std::string file = reload_file;
  if (reload_type == "same") {
    // GrainsExec::restartFileName_AorB: the last entry of the table
    std::ifstream rft(save_root + "_RFTable.txt");
    std::string tt, rf;
    file = save_root;
    while (rft >> tt >> rf) file = rf;
  }
  std::ifstream probe(file);
  if (!probe && !file.empty() && file[0] != '/') file = dir + file;
  if ((e = grains_gpu_load_reload(s, file.c_str()))) return e;
  if (reload_velocity == "Reset") {
    // AllComponents::resetKinematics -> ParticleKinematics::reset
    const size_t n = (size_t)s->nparticles;
    std::vector<double> z(3 * n, 0.), q(4 * n, 0.);
    for (size_t i = 0; i < n; ++i) q[4 * i + 3] = 1.;
    if ((e = grains_gpu_set_state(s, nullptr, z.data(), z.data(), q.data(), 0))) return e;
  }
  // the reload's time is the start of the run (Grains.cpp:822-826)
  return GRAINS_OK;
}
