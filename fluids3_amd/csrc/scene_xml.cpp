// scene_xml.cpp -- scene files in the reference's scene.xml schema -> fx_scene.
//
// The schema (the reference's README.md "Scene parameters"; the example file
// fluids5.0/assets/scene.xml) is one <Fluid> element followed by
// <Scene id="k"> elements, each a flat list of <Key> value </Key> parameters.
// Any parameter may appear in either section and a scene's values override
// the Fluid section's: the reference parses "Fluid" and then the chosen
// scene into the same parameter set (FluidSystem::ParseXML,
// fluids3.2/fluids/fluid_system.cpp:2268-2309).  Vectors are written
// "x, y, z".  The key table keeps the reference's names; keys of forces the
// 5.0 step does not have (point gravity, the X+ wave wall) and of the viewer
// are returned to the caller as unused, not dropped silently.
//
// The parser reads exactly that shape: top-level elements (comments, the XML
// declaration and processing instructions skipped), each holding text-only
// child elements; the five predefined entities are decoded in values.
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "fx_common.hpp"
#include "fx_scene.hpp"

namespace {

struct Section {
  std::string tag, id;
  std::vector<std::pair<std::string, std::string>> keys;  // in file order
};

class Reader {
 public:
  explicit Reader(const char* text) : p_(text), e_(text + strlen(text)) {}

  // all top-level sections; false with msg set on malformed text
  bool sections(std::vector<Section>& out, std::string& msg) {
    while (true) {
      if (!skip_misc(msg)) return false;
      if (p_ == e_) return true;
      if (*p_ != '<') return fail(msg, "text outside an element");
      Section s;
      bool empty = false;
      std::vector<std::pair<std::string, std::string>> attrs;
      if (!open_tag(s.tag, attrs, empty, msg)) return false;
      for (auto& a : attrs)
        if (a.first == "id") s.id = trim(a.second);
      if (!empty && !children(s, msg)) return false;
      out.push_back(std::move(s));
    }
  }

 private:
  const char* p_;
  const char* e_;

  bool fail(std::string& msg, const char* what) {
    msg = what;
    return false;
  }
  static bool name_char(char c) {
    return (c >= 'A' && c <= 'Z') || (c >= 'a' && c <= 'z') || (c >= '0' && c <= '9') || c == '_' || c == '-' ||
           c == '.' || c == ':';
  }
  static bool space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; }
  static std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && space(s[a])) a++;
    while (b > a && space(s[b - 1])) b--;
    return s.substr(a, b - a);
  }
  bool starts(const char* lit) const {
    const size_t n = strlen(lit);
    return (size_t)(e_ - p_) >= n && memcmp(p_, lit, n) == 0;
  }
  bool skip_to(const char* lit) {  // past the next occurrence of lit
    const size_t n = strlen(lit);
    for (; (size_t)(e_ - p_) >= n; p_++)
      if (memcmp(p_, lit, n) == 0) {
        p_ += n;
        return true;
      }
    p_ = e_;
    return false;
  }
  // whitespace, comments, <?...?> and <!DOCTYPE ...>
  bool skip_misc(std::string& msg) {
    while (true) {
      while (p_ < e_ && space(*p_)) p_++;
      if (starts("<!--")) {
        if (!skip_to("-->")) return fail(msg, "unterminated comment");
      } else if (starts("<?")) {
        if (!skip_to("?>")) return fail(msg, "unterminated processing instruction");
      } else if (starts("<!")) {
        if (!skip_to(">")) return fail(msg, "unterminated declaration");
      } else {
        return true;
      }
    }
  }
  bool name(std::string& out, std::string& msg) {
    const char* a = p_;
    while (p_ < e_ && name_char(*p_)) p_++;
    if (p_ == a) return fail(msg, "missing element or attribute name");
    out.assign(a, p_);
    return true;
  }
  // at '<': reads <name attr="v" ...> or <name .../>
  bool open_tag(std::string& tag, std::vector<std::pair<std::string, std::string>>& attrs, bool& empty,
                std::string& msg) {
    p_++;
    if (!name(tag, msg)) return false;
    while (true) {
      while (p_ < e_ && space(*p_)) p_++;
      if (p_ == e_) return fail(msg, "unterminated start tag");
      if (*p_ == '>') {
        p_++;
        empty = false;
        return true;
      }
      if (starts("/>")) {
        p_ += 2;
        empty = true;
        return true;
      }
      std::string an;
      if (!name(an, msg)) return false;
      while (p_ < e_ && space(*p_)) p_++;
      if (p_ == e_ || *p_ != '=') return fail(msg, "attribute without a value");
      p_++;
      while (p_ < e_ && space(*p_)) p_++;
      if (p_ == e_ || (*p_ != '"' && *p_ != '\'')) return fail(msg, "unquoted attribute value");
      const char q = *p_++;
      const char* a = p_;
      while (p_ < e_ && *p_ != q) p_++;
      if (p_ == e_) return fail(msg, "unterminated attribute value");
      std::string v;
      if (!decode(a, p_, v, msg)) return false;
      p_++;
      attrs.emplace_back(an, v);
    }
  }
  bool close_tag(const std::string& tag, std::string& msg) {
    if (!starts("</")) return fail(msg, "missing end tag");
    p_ += 2;
    std::string n;
    if (!name(n, msg)) return false;
    while (p_ < e_ && space(*p_)) p_++;
    if (p_ == e_ || *p_ != '>' || n != tag) {
      msg = "end tag </" + n + "> does not close <" + tag + ">";
      return false;
    }
    p_++;
    return true;
  }
  bool decode(const char* a, const char* b, std::string& out, std::string& msg) {
    static const struct {
      const char* ent;
      char c;
    } kEnt[] = {{"&lt;", '<'}, {"&gt;", '>'}, {"&amp;", '&'}, {"&quot;", '"'}, {"&apos;", '\''}};
    out.clear();
    while (a < b) {
      if (*a == '&') {
        bool ok = false;
        for (const auto& e : kEnt) {
          const size_t n = strlen(e.ent);
          if ((size_t)(b - a) >= n && memcmp(a, e.ent, n) == 0) {
            out.push_back(e.c);
            a += n;
            ok = true;
            break;
          }
        }
        if (!ok) return fail(msg, "unknown entity");
      } else {
        out.push_back(*a++);
      }
    }
    return true;
  }
  // the children of a section: text-only elements, up to the section's end tag
  bool children(Section& s, std::string& msg) {
    while (true) {
      if (!skip_misc(msg)) return false;
      if (p_ == e_) return fail(msg, "unterminated section");
      if (starts("</")) return close_tag(s.tag, msg);
      if (*p_ != '<') return fail(msg, "text between parameters");
      std::string key;
      std::vector<std::pair<std::string, std::string>> attrs;
      bool empty = false;
      if (!open_tag(key, attrs, empty, msg)) return false;
      std::string val;
      if (!empty) {
        const char* a = p_;
        while (p_ < e_ && *p_ != '<') p_++;
        if (!decode(a, p_, val, msg)) return false;
        if (!skip_misc(msg)) return false;  // a comment after the value
        if (!close_tag(key, msg)) return false;
      }
      s.keys.emplace_back(key, trim(val));
    }
  }
};

bool number(const std::string& s, double& v) {
  if (s.empty()) return false;
  errno = 0;
  char* end = nullptr;
  v = strtod(s.c_str(), &end);
  return errno == 0 && end == s.c_str() + s.size();
}

bool vector3(const std::string& s, fx_float3& v) {
  std::string t = s;
  for (char& c : t)
    if (c == ',') c = ' ';
  double d[4];
  int k = 0;
  const char* p = t.c_str();
  while (*p) {
    while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r') p++;
    if (!*p) break;
    if (k == 3) return false;
    char* end = nullptr;
    errno = 0;
    d[k] = strtod(p, &end);
    if (end == p || errno != 0 || (*end && *end != ' ' && *end != '\t' && *end != '\n' && *end != '\r')) return false;
    k++;
    p = end;
  }
  if (k != 3) return false;
  v = fx_float3{(float)d[0], (float)d[1], (float)d[2]};
  return true;
}

struct ScalarKey {
  const char* key;
  float fx_scene::*field;
};
const ScalarKey kScalar[] = {
    {"DT", &fx_scene::dt},
    {"SimScale", &fx_scene::sim_scale},
    {"Viscosity", &fx_scene::visc},
    {"RestDensity", &fx_scene::rest_density},
    {"Mass", &fx_scene::mass},
    {"Radius", &fx_scene::pradius},
    {"IntStiff", &fx_scene::int_stiff},
    {"BoundStiff", &fx_scene::bound_stiff},
    {"BoundDamp", &fx_scene::bound_damp},
    {"AccelLimit", &fx_scene::accel_limit},
    {"VelLimit", &fx_scene::vel_limit},
    {"GroundSlope", &fx_scene::bound_slope},
    {"WaveForceMin", &fx_scene::bound_wall_force},  // amplitude of the X- wave wall
    {"WaveForceFreq", &fx_scene::bound_wall_freq},
    {"GridDensity", &fx_scene::grid_density},
    {"Spacing", &fx_scene::spacing},
    {"SmoothRadius", &fx_scene::smooth_radius},
};
struct VectorKey {
  const char* key;
  fx_float3 fx_scene::*field;
};
const VectorKey kVector[] = {{"VolMin", &fx_scene::bound_min},
                             {"VolMax", &fx_scene::bound_max},
                             {"InitMin", &fx_scene::init_min},
                             {"InitMax", &fx_scene::init_max}};
// schema keys the step does not use: point gravity, the X+ wave wall, viewer settings
const char* const kUnused[] = {"PointGravAmt", "PointGravPos", "WaveForceMax", "SearchDist",
                               "DrawMode",     "DrawGrid",     "DrawText"};

int parse(const char* text, std::vector<Section>& secs, const char* who) {
  if (text == nullptr) return fx::fail(FX_EINVAL, "%s: null text", who);
  std::string msg;
  Reader r(text);
  if (!r.sections(secs, msg)) return fx::fail(FX_EINVAL, "%s: not a scene file: %s", who, msg.c_str());
  return FX_OK;
}

}  // namespace

extern "C" int fx_scene_xml_ids(const char* text, int32_t* ids, int32_t max_ids, int32_t* count) {
  if (count == nullptr || (ids == nullptr && max_ids > 0)) return fx::fail(FX_EINVAL, "fx_scene_xml_ids: null argument");
  std::vector<Section> secs;
  FX_TRY(parse(text, secs, "fx_scene_xml_ids"));
  int32_t k = 0;
  for (const Section& s : secs) {
    if (s.tag != "Scene") continue;
    double v;
    if (!number(s.id, v) || v != (double)(int32_t)v)
      return fx::fail(FX_EINVAL, "fx_scene_xml_ids: <Scene id=\"%s\"> is not an integer", s.id.c_str());
    if (k < max_ids) ids[k] = (int32_t)v;
    k++;
  }
  *count = k;
  return k > max_ids ? fx::fail(FX_ERANGE, "fx_scene_xml_ids: %d scenes, room for %d", k, max_ids) : FX_OK;
}

extern "C" int fx_scene_from_xml(const char* text, int32_t scene_id, fx_scene* scene, char* name, int32_t name_len,
                                 char* unused, int32_t unused_len) {
  if (scene == nullptr) return fx::fail(FX_EINVAL, "fx_scene_from_xml: null scene");
  std::vector<Section> secs;
  FX_TRY(parse(text, secs, "fx_scene_from_xml"));
  const Section* fluid = nullptr;
  const Section* chosen = nullptr;
  for (const Section& s : secs) {
    if (s.tag == "Fluid") fluid = &s;
    double v;
    if (s.tag == "Scene" && number(s.id, v) && v == (double)scene_id) chosen = &s;  // the last of an id wins
  }
  if (chosen == nullptr) return fx::fail(FX_EINVAL, "fx_scene_from_xml: no <Scene id=\"%d\">", scene_id);
  fx_scene sc = *scene;  // the caller's values are the defaults
  std::string nm, un;
  for (const Section* sec : {fluid, chosen}) {
    if (sec == nullptr) continue;
    for (const auto& kv : sec->keys) {
      const std::string& k = kv.first;
      const std::string& v = kv.second;
      bool done = false;
      double d;
      if (k == "Name") {
        nm = v;
        done = true;
      } else if (k == "Num") {
        if (!number(v, d) || d < 0 || d > 2147483647.0)
          return fx::fail(FX_EINVAL, "fx_scene_from_xml: <Num> %s is not a particle count", v.c_str());
        sc.num_points = (int32_t)d;
        done = true;
      } else if (k == "PlaneGravDir") {
        if (!vector3(v, sc.grav_dir))
          return fx::fail(FX_EINVAL, "fx_scene_from_xml: <%s> expects three numbers, got \"%s\"", k.c_str(), v.c_str());
        sc.grav_amt = 1.0f;
        done = true;
      }
      for (const ScalarKey& s : kScalar)
        if (!done && k == s.key) {
          if (!number(v, d))
            return fx::fail(FX_EINVAL, "fx_scene_from_xml: <%s> expects a number, got \"%s\"", k.c_str(), v.c_str());
          sc.*s.field = (float)d;
          done = true;
        }
      for (const VectorKey& s : kVector)
        if (!done && k == s.key) {
          if (!vector3(v, sc.*s.field))
            return fx::fail(FX_EINVAL, "fx_scene_from_xml: <%s> expects three numbers, got \"%s\"", k.c_str(),
                            v.c_str());
          done = true;
        }
      for (const char* u : kUnused)
        if (!done && k == u) {
          un += k + "=" + v + "\n";
          done = true;
        }
      if (!done) return fx::fail(FX_EINVAL, "fx_scene_from_xml: unknown key <%s>", k.c_str());
    }
  }
  if (sc.num_points <= 0) return fx::fail(FX_EINVAL, "fx_scene_from_xml: scene %d: <Num> missing or not positive", scene_id);
  if (name != nullptr && name_len > 0) {
    const size_t n = nm.size() < (size_t)name_len - 1 ? nm.size() : (size_t)name_len - 1;
    memcpy(name, nm.data(), n);
    name[n] = '\0';
  }
  if (unused != nullptr && unused_len > 0) {
    if (un.size() + 1 > (size_t)unused_len)
      return fx::fail(FX_ERANGE, "fx_scene_from_xml: unused keys need %zu bytes, room for %d", un.size() + 1, unused_len);
    memcpy(unused, un.c_str(), un.size() + 1);
  }
  *scene = sc;
  return FX_OK;
}
