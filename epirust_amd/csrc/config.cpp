// Minimal JSON reader + the reference's Config schema.
//
// Follows serde's behaviour for the standalone Config (common/src/config/mod.rs:44-58):
// unknown keys are ignored, `disease_overrides` / `output_file` are accepted and unused,
// u32 fields must be non-negative integer literals, f64 fields accept any number, the
// `population` and each `interventions` entry are externally tagged enums.
#include "config.hpp"

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>

namespace epi {
namespace {

struct JValue {
  enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
  bool b = false;
  double num = 0;
  bool is_int = false;   // literal had no fraction / exponent
  bool negative = false;
  uint64_t uval = 0;     // exact value of a non-negative integer literal
  std::string str;
  std::vector<JValue> arr;
  std::vector<std::pair<std::string, JValue>> obj;

  const JValue* get(const std::string& k) const {
    for (auto& kv : obj)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
};

class Reader {
 public:
  explicit Reader(const std::string& s) : s_(s) {}
  bool parse(JValue& v, std::string& err) {
    if (!value(v, err)) return false;
    ws();
    if (i_ != s_.size()) return fail(err, "trailing characters");
    return true;
  }

 private:
  const std::string& s_;
  size_t i_ = 0;
  int depth_ = 0;

  bool fail(std::string& err, const char* what) {
    size_t line = 1, col = 1;
    for (size_t k = 0; k < i_ && k < s_.size(); ++k) {
      if (s_[k] == '\n') { ++line; col = 1; } else { ++col; }
    }
    err = std::string(what) + " at line " + std::to_string(line) + " column " + std::to_string(col);
    return false;
  }
  void ws() {
    while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) ++i_;
  }
  bool lit(const char* w) {
    size_t n = std::char_traits<char>::length(w);
    if (s_.compare(i_, n, w) == 0) { i_ += n; return true; }
    return false;
  }
  bool value(JValue& v, std::string& err) {
    ws();
    if (i_ >= s_.size()) return fail(err, "EOF while parsing a value");
    if (++depth_ > 128) return fail(err, "recursion limit exceeded");
    bool ok = true;
    char c = s_[i_];
    if (c == '{') ok = object(v, err);
    else if (c == '[') ok = array(v, err);
    else if (c == '"') { v.kind = JValue::Str; ok = string(v.str, err); }
    else if (lit("true")) { v.kind = JValue::Bool; v.b = true; }
    else if (lit("false")) { v.kind = JValue::Bool; v.b = false; }
    else if (lit("null")) { v.kind = JValue::Null; }
    else if (c == '-' || (c >= '0' && c <= '9')) ok = number(v, err);
    else ok = fail(err, "expected value");
    --depth_;
    return ok;
  }
  bool object(JValue& v, std::string& err) {
    v.kind = JValue::Obj;
    ++i_;
    ws();
    if (i_ < s_.size() && s_[i_] == '}') { ++i_; return true; }
    for (;;) {
      ws();
      if (i_ >= s_.size() || s_[i_] != '"') return fail(err, "key must be a string");
      std::string k;
      if (!string(k, err)) return false;
      ws();
      if (i_ >= s_.size() || s_[i_] != ':') return fail(err, "expected `:`");
      ++i_;
      JValue child;
      if (!value(child, err)) return false;
      v.obj.emplace_back(std::move(k), std::move(child));
      ws();
      if (i_ < s_.size() && s_[i_] == ',') { ++i_; continue; }
      if (i_ < s_.size() && s_[i_] == '}') { ++i_; return true; }
      return fail(err, "expected `,` or `}`");
    }
  }
  bool array(JValue& v, std::string& err) {
    v.kind = JValue::Arr;
    ++i_;
    ws();
    if (i_ < s_.size() && s_[i_] == ']') { ++i_; return true; }
    for (;;) {
      JValue child;
      if (!value(child, err)) return false;
      v.arr.push_back(std::move(child));
      ws();
      if (i_ < s_.size() && s_[i_] == ',') { ++i_; continue; }
      if (i_ < s_.size() && s_[i_] == ']') { ++i_; return true; }
      return fail(err, "expected `,` or `]`");
    }
  }
  static void put_utf8(std::string& o, uint32_t cp) {
    if (cp < 0x80) o += char(cp);
    else if (cp < 0x800) { o += char(0xC0 | (cp >> 6)); o += char(0x80 | (cp & 0x3F)); }
    else if (cp < 0x10000) { o += char(0xE0 | (cp >> 12)); o += char(0x80 | ((cp >> 6) & 0x3F)); o += char(0x80 | (cp & 0x3F)); }
    else { o += char(0xF0 | (cp >> 18)); o += char(0x80 | ((cp >> 12) & 0x3F)); o += char(0x80 | ((cp >> 6) & 0x3F)); o += char(0x80 | (cp & 0x3F)); }
  }
  bool hex4(uint32_t& cp, std::string& err) {
    if (i_ + 4 > s_.size()) return fail(err, "invalid escape");
    cp = 0;
    for (int k = 0; k < 4; ++k) {
      char h = s_[i_++];
      cp <<= 4;
      if (h >= '0' && h <= '9') cp |= uint32_t(h - '0');
      else if (h >= 'a' && h <= 'f') cp |= uint32_t(h - 'a' + 10);
      else if (h >= 'A' && h <= 'F') cp |= uint32_t(h - 'A' + 10);
      else return fail(err, "invalid escape");
    }
    return true;
  }
  bool string(std::string& o, std::string& err) {
    ++i_;
    while (i_ < s_.size()) {
      char c = s_[i_++];
      if (c == '"') return true;
      if (static_cast<unsigned char>(c) < 0x20) return fail(err, "control character in string");
      if (c != '\\') { o += c; continue; }
      if (i_ >= s_.size()) break;
      char e = s_[i_++];
      switch (e) {
        case '"': o += '"'; break;
        case '\\': o += '\\'; break;
        case '/': o += '/'; break;
        case 'b': o += '\b'; break;
        case 'f': o += '\f'; break;
        case 'n': o += '\n'; break;
        case 'r': o += '\r'; break;
        case 't': o += '\t'; break;
        case 'u': {
          uint32_t cp = 0;
          if (!hex4(cp, err)) return false;
          if (cp >= 0xD800 && cp < 0xDC00 && s_.compare(i_, 2, "\\u") == 0) {
            i_ += 2;
            uint32_t lo = 0;
            if (!hex4(lo, err)) return false;
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          }
          put_utf8(o, cp);
          break;
        }
        default: return fail(err, "invalid escape");
      }
    }
    return fail(err, "EOF while parsing a string");
  }
  bool number(JValue& v, std::string& err) {
    size_t start = i_;
    v.kind = JValue::Num;
    if (s_[i_] == '-') { v.negative = true; ++i_; }
    if (i_ >= s_.size() || !(s_[i_] >= '0' && s_[i_] <= '9')) return fail(err, "invalid number");
    if (s_[i_] == '0') ++i_;
    else while (i_ < s_.size() && s_[i_] >= '0' && s_[i_] <= '9') ++i_;
    bool frac = false;
    if (i_ < s_.size() && s_[i_] == '.') {
      frac = true;
      ++i_;
      if (i_ >= s_.size() || !(s_[i_] >= '0' && s_[i_] <= '9')) return fail(err, "invalid number");
      while (i_ < s_.size() && s_[i_] >= '0' && s_[i_] <= '9') ++i_;
    }
    if (i_ < s_.size() && (s_[i_] == 'e' || s_[i_] == 'E')) {
      frac = true;
      ++i_;
      if (i_ < s_.size() && (s_[i_] == '+' || s_[i_] == '-')) ++i_;
      if (i_ >= s_.size() || !(s_[i_] >= '0' && s_[i_] <= '9')) return fail(err, "invalid number");
      while (i_ < s_.size() && s_[i_] >= '0' && s_[i_] <= '9') ++i_;
    }
    std::string tok = s_.substr(start, i_ - start);
    v.num = std::strtod(tok.c_str(), nullptr);  // correctly rounded, like serde_json
    v.is_int = !frac;
    if (v.is_int && !v.negative) {
      errno = 0;
      unsigned long long u = std::strtoull(tok.c_str(), nullptr, 10);
      if (errno == ERANGE) v.is_int = false;  // too large for an integer field
      v.uval = u;
    }
    return true;
  }
};

struct Ctx {
  std::string err;
  bool fail(const std::string& e) {
    if (err.empty()) err = e;
    return false;
  }
};

const JValue* need(Ctx& c, const JValue& o, const char* key) {
  const JValue* v = o.get(key);
  if (!v) c.fail(std::string("missing field `") + key + "`");
  return v;
}

bool get_u32(Ctx& c, const JValue& o, const char* key, uint32_t& out) {
  const JValue* v = need(c, o, key);
  if (!v) return false;
  if (v->kind != JValue::Num) return c.fail(std::string("invalid type for `") + key + "`, expected u32");
  if (!v->is_int || v->negative) {
    if (v->negative && v->is_int) return c.fail(std::string("invalid value: integer `") + key + "`, expected u32");
    return c.fail(std::string("invalid type: floating point for `") + key + "`, expected u32");
  }
  if (v->uval > 0xFFFFFFFFull) return c.fail(std::string("invalid value for `") + key + "`, expected u32");
  out = static_cast<uint32_t>(v->uval);
  return true;
}

bool get_f64(Ctx& c, const JValue& o, const char* key, double& out) {
  const JValue* v = need(c, o, key);
  if (!v) return false;
  if (v->kind != JValue::Num) return c.fail(std::string("invalid type for `") + key + "`, expected f64");
  out = v->num;
  return true;
}

bool tagged(Ctx& c, const JValue& v, const char* what, std::string& tag, const JValue*& body) {
  if (v.kind == JValue::Str) {
    tag = v.str;
    body = nullptr;
    return true;
  }
  if (v.kind != JValue::Obj || v.obj.size() != 1) return c.fail(std::string("invalid type for ") + what + ", expected an enum");
  tag = v.obj[0].first;
  body = &v.obj[0].second;
  return true;
}

}  // namespace

bool parse_config(const std::string& text, Config& cfg, std::string& err) {
  JValue root;
  Reader rd(text);
  if (!rd.parse(root, err)) return false;
  Ctx c;
  if (root.kind != JValue::Obj) { err = "invalid type: expected struct Config"; return false; }

  // population
  if (const JValue* pop = need(c, root, "population")) {
    std::string tag;
    const JValue* body = nullptr;
    if (tagged(c, *pop, "population", tag, body)) {
      if (tag == "Auto" && body && body->kind == JValue::Obj) {
        get_u32(c, *body, "number_of_agents", cfg.number_of_agents);
        get_f64(c, *body, "public_transport_percentage", cfg.public_transport_percentage);
        get_f64(c, *body, "working_percentage", cfg.working_percentage);
      } else if (tag == "Csv" && body && body->kind == JValue::Obj) {
        const JValue* file = need(c, *body, "file");
        const JValue* cols = need(c, *body, "cols");
        if (file && file->kind != JValue::Str) c.fail("invalid type for `file`, expected a string");
        else if (cols && cols->kind != JValue::Arr) c.fail("invalid type for `cols`, expected a sequence");
        else if (file) {
          cfg.csv_population = true;
          cfg.csv_file = file->str;
        }
      } else {
        c.fail("unknown variant `" + tag + "`, expected `Csv` or `Auto`");
      }
    }
  }
  // disease: Option<Disease>; a missing disease is the reference's unwrap() panic
  const JValue* d = root.get("disease");
  if (!d || d->kind == JValue::Null) {
    c.fail("called `Option::unwrap()` on a `None` value (config has no disease)");
  } else if (d->kind != JValue::Obj) {
    c.fail("invalid type for `disease`, expected struct Disease");
  } else {
    get_u32(c, *d, "regular_transmission_start_day", cfg.regular_transmission_start_day);
    get_u32(c, *d, "high_transmission_start_day", cfg.high_transmission_start_day);
    get_u32(c, *d, "last_day", cfg.last_day);
    get_u32(c, *d, "asymptomatic_last_day", cfg.asymptomatic_last_day);
    get_u32(c, *d, "mild_infected_last_day", cfg.mild_infected_last_day);
    get_f64(c, *d, "regular_transmission_rate", cfg.regular_transmission_rate);
    get_f64(c, *d, "high_transmission_rate", cfg.high_transmission_rate);
    get_f64(c, *d, "death_rate", cfg.death_rate);
    get_f64(c, *d, "percentage_asymptomatic_population", cfg.percentage_asymptomatic_population);
    get_f64(c, *d, "percentage_severe_infected_population", cfg.percentage_severe_infected_population);
    get_u32(c, *d, "exposed_duration", cfg.exposed_duration);
    get_u32(c, *d, "pre_symptomatic_duration", cfg.pre_symptomatic_duration);
  }
  if (const JValue* geo = need(c, root, "geography_parameters")) {
    if (geo->kind != JValue::Obj) c.fail("invalid type for `geography_parameters`");
    else {
      get_u32(c, *geo, "grid_size", cfg.grid_size);
      get_f64(c, *geo, "hospital_beds_percentage", cfg.hospital_beds_percentage);
    }
  }
  get_u32(c, root, "hours", cfg.hours);
  if (const JValue* ivs = need(c, root, "interventions")) {
    if (ivs->kind != JValue::Arr) c.fail("invalid type for `interventions`, expected a sequence");
    else
      for (const JValue& iv : ivs->arr) {
        std::string tag;
        const JValue* body = nullptr;
        if (!tagged(c, iv, "intervention", tag, body)) break;
        if (!body || body->kind != JValue::Obj) { c.fail("invalid type for intervention `" + tag + "`"); break; }
        if (tag == "Vaccinate") {
          VaccinateConfig v{};
          get_u32(c, *body, "at_hour", v.at_hour);
          get_f64(c, *body, "percent", v.percent);
          cfg.vaccinate.push_back(v);
        } else if (tag == "Lockdown") {
          uint32_t at = 0;
          double ess = 0;
          get_u32(c, *body, "at_number_of_infections", at);
          get_f64(c, *body, "essential_workers_population", ess);
          if (!cfg.has_lockdown) {
            cfg.has_lockdown = true;
            cfg.lockdown_at_number_of_infections = at;
            cfg.essential_workers_population = ess;
          }
        } else if (tag == "BuildNewHospital") {
          uint32_t thr = 0;
          get_u32(c, *body, "spread_rate_threshold", thr);
          if (!cfg.has_build_new_hospital) {
            cfg.has_build_new_hospital = true;
            cfg.spread_rate_threshold = thr;
          }
        } else {
          c.fail("unknown variant `" + tag + "`, expected one of `Vaccinate`, `Lockdown`, `BuildNewHospital`");
        }
      }
  }
  if (const JValue* e = root.get("enable_citizen_state_messages")) {
    if (e->kind != JValue::Bool) c.fail("invalid type for `enable_citizen_state_messages`, expected a boolean");
    else cfg.enable_citizen_state_messages = e->b;
  }
  if (const JValue* si = root.get("starting_infections")) {
    if (si->kind != JValue::Obj) c.fail("invalid type for `starting_infections`");
    else {
      get_u32(c, *si, "infected_mild_asymptomatic", cfg.si_asymptomatic);
      get_u32(c, *si, "infected_mild_symptomatic", cfg.si_mild);
      get_u32(c, *si, "infected_severe", cfg.si_severe);
      get_u32(c, *si, "exposed", cfg.si_exposed);
    }
  }
  if (const JValue* sd = root.get("seed")) {
    if (sd->kind != JValue::Num || !sd->is_int || sd->negative) c.fail("invalid type for `seed`, expected u64");
    else { cfg.seed = sd->uval; cfg.has_seed = true; }
  }
  if (!c.err.empty()) { err = c.err; return false; }
  return true;
}

namespace {

// One CSV record split on commas; a field wrapped in double quotes may contain commas and
// doubled quotes.
std::vector<std::string> split_csv(const std::string& line) {
  std::vector<std::string> out;
  std::string cur;
  bool quoted = false;
  for (size_t i = 0; i < line.size(); ++i) {
    const char ch = line[i];
    if (quoted) {
      if (ch == '"' && i + 1 < line.size() && line[i + 1] == '"') { cur += '"'; ++i; }
      else if (ch == '"') quoted = false;
      else cur += ch;
    } else if (ch == '"') {
      quoted = true;
    } else if (ch == ',') {
      out.push_back(cur);
      cur.clear();
    } else if (ch != '\r') {
      cur += ch;
    }
  }
  out.push_back(cur);
  return out;
}

}  // namespace

bool read_population_csv(Config& cfg, std::string& err) {
  std::ifstream f(cfg.csv_file);
  if (!f) {
    err = "Could not read population file: " + cfg.csv_file;
    return false;
  }
  std::string line;
  if (!std::getline(f, line)) {
    cfg.number_of_agents = 0;
    return true;
  }
  const std::vector<std::string> head = split_csv(line);
  int col_ind = -1, col_age = -1, col_working = -1, col_pt = -1;
  for (size_t k = 0; k < head.size(); ++k) {
    if (head[k] == "ind") col_ind = static_cast<int>(k);
    else if (head[k] == "age") col_age = static_cast<int>(k);
    else if (head[k] == "working") col_working = static_cast<int>(k);
    else if (head[k] == "pub_transport") col_pt = static_cast<int>(k);
  }
  const char* missing = col_ind < 0 ? "ind" : col_age < 0 ? "age" : col_working < 0 ? "working"
                        : col_pt < 0 ? "pub_transport" : nullptr;
  if (missing) {
    err = std::string("Could not deserialize population line: missing field `") + missing + "`";
    return false;
  }
  cfg.csv_working.clear();
  cfg.csv_pub_transport.clear();
  size_t row = 1;
  auto boolean = [&](const std::string& v, uint8_t& out) {
    if (v == "True") { out = 1; return true; }
    if (v == "False") { out = 0; return true; }
    err = "Could not deserialize population line " + std::to_string(row) + ": invalid value: string \"" + v +
          "\", expected True or False";
    return false;
  };
  while (std::getline(f, line)) {
    ++row;
    if (line.empty() || line == "\r") continue;
    const std::vector<std::string> v = split_csv(line);
    if (static_cast<int>(v.size()) <= std::max(std::max(col_ind, col_age), std::max(col_working, col_pt))) {
      err = "Could not deserialize population line " + std::to_string(row) + ": too few fields";
      return false;
    }
    const std::string& ind = v[col_ind];
    if (ind.empty() || ind.find_first_not_of("0123456789") != std::string::npos || ind.size() > 10 ||
        std::stoull(ind) > 0xFFFFFFFFull) {
      err = "Could not deserialize population line " + std::to_string(row) + ": invalid `ind` \"" + ind + "\"";
      return false;
    }
    uint8_t w = 0, pt = 0;
    if (!boolean(v[col_working], w) || !boolean(v[col_pt], pt)) return false;
    cfg.csv_working.push_back(w);
    cfg.csv_pub_transport.push_back(pt);
  }
  cfg.number_of_agents = static_cast<uint32_t>(cfg.csv_working.size());
  return true;
}

}  // namespace epi
