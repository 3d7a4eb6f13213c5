// proto_io.cpp -- protobuf wire encoding of the snapshot messages and the
// record framing (declarations and conventions in proto_io.h).
#include "proto_io.h"

#include <cstring>

namespace UVLM {
namespace proto {
namespace {

// Wire types of the protobuf encoding.
constexpr int kVarint = 0, kFixed64 = 1, kLength = 2, kFixed32 = 5;

void put_varint(std::string* out, uint64_t v) {
  while (v >= 0x80) {
    out->push_back(static_cast<char>((v & 0x7f) | 0x80));
    v >>= 7;
  }
  out->push_back(static_cast<char>(v));
}
void put_key(std::string* out, int field, int wire) {
  put_varint(out, (static_cast<uint64_t>(field) << 3) | static_cast<uint64_t>(wire));
}
void put_double(std::string* out, int field, double v) {
  put_key(out, field, kFixed64);
  uint64_t bits;
  std::memcpy(&bits, &v, sizeof(bits));
  for (int k = 0; k < 8; ++k) out->push_back(static_cast<char>((bits >> (8 * k)) & 0xff));
}
void put_uint(std::string* out, int field, uint64_t v) {
  put_key(out, field, kVarint);
  put_varint(out, v);
}
void put_bytes(std::string* out, int field, const std::string& b) {
  put_key(out, field, kLength);
  put_varint(out, b.size());
  out->append(b);
}

// message Point: x = 1, y = 2, z = 3 (all set by Vector3dToPoint)
std::string enc(const Point& p) {
  std::string s;
  put_double(&s, 1, p.x());
  put_double(&s, 2, p.y());
  put_double(&s, 3, p.z());
  return s;
}
std::string enc(const VortexRing& v) {
  std::string s;
  if (v.has_gamma) put_double(&s, 1, v.gamma);
  for (const auto& p : v.nodes) put_bytes(&s, 2, enc(p));
  for (const auto& p : v.nodes0) put_bytes(&s, 3, enc(p));
  return s;
}
std::string enc(const MorphingVelocity& m) {
  std::string s;
  if (m.has_center) put_bytes(&s, 1, enc(m.center));
  for (const auto& p : m.nodes) put_bytes(&s, 2, enc(p));
  if (m.has_freestream) put_bytes(&s, 3, enc(m.freestream));
  return s;
}
std::string enc(const ContainerShape& c) {
  std::string s;
  put_uint(&s, 1, c.rows);
  put_uint(&s, 2, c.cols);
  if (c.has_id) put_uint(&s, 3, c.id);
  if (c.has_origin) put_bytes(&s, 4, enc(c.origin));
  if (c.has_chord) put_double(&s, 5, c.chord);
  if (c.has_span) put_double(&s, 6, c.span);
  return s;
}

// --- parsing -----------------------------------------------------------------
struct Cursor {
  const unsigned char* p;
  const unsigned char* end;
  bool ok = true;
};

bool get_varint(Cursor* c, uint64_t* v) {
  uint64_t r = 0;
  for (int shift = 0; shift < 64; shift += 7) {
    if (c->p >= c->end) return c->ok = false;
    const unsigned char b = *c->p++;
    r |= static_cast<uint64_t>(b & 0x7f) << shift;
    if (!(b & 0x80)) {
      *v = r;
      return true;
    }
  }
  return c->ok = false;
}
bool get_double(Cursor* c, double* v) {
  if (c->end - c->p < 8) return c->ok = false;
  uint64_t bits = 0;
  for (int k = 0; k < 8; ++k) bits |= static_cast<uint64_t>(c->p[k]) << (8 * k);
  c->p += 8;
  std::memcpy(v, &bits, sizeof(*v));
  return true;
}
bool get_bytes(Cursor* c, Cursor* sub) {
  uint64_t n;
  if (!get_varint(c, &n)) return false;
  if (static_cast<uint64_t>(c->end - c->p) < n) return c->ok = false;
  sub->p = c->p;
  sub->end = c->p + n;
  sub->ok = true;
  c->p += n;
  return true;
}
bool skip(Cursor* c, int wire) {
  uint64_t v;
  Cursor sub{};
  switch (wire) {
    case kVarint:
      return get_varint(c, &v);
    case kFixed64:
      if (c->end - c->p < 8) return c->ok = false;
      c->p += 8;
      return true;
    case kLength:
      return get_bytes(c, &sub);
    case kFixed32:
      if (c->end - c->p < 4) return c->ok = false;
      c->p += 4;
      return true;
    default:
      return c->ok = false;
  }
}

// Calls f(field, wire, cursor) for every field; f consumes the value or
// returns false to have it skipped.
template <class F>
bool fields(Cursor c, F&& f) {
  while (c.p < c.end) {
    uint64_t key;
    if (!get_varint(&c, &key)) return false;
    const int field = static_cast<int>(key >> 3), wire = static_cast<int>(key & 7);
    if (!f(field, wire, &c) && !skip(&c, wire)) return false;
    if (!c.ok) return false;
  }
  return c.ok;
}

bool dec(Cursor c, Point* p) {
  return fields(c, [&](int f, int w, Cursor* k) {
    double v;
    if (w != kFixed64 || f < 1 || f > 3) return false;
    if (!get_double(k, &v)) return true;
    if (f == 1) p->set_x(v);
    if (f == 2) p->set_y(v);
    if (f == 3) p->set_z(v);
    return true;
  });
}
bool dec(Cursor c, VortexRing* v) {
  bool ok = true;
  const bool r = fields(c, [&](int f, int w, Cursor* k) {
    Cursor sub{};
    if (f == 1 && w == kFixed64) {
      v->has_gamma = get_double(k, &v->gamma);
      return true;
    }
    if ((f == 2 || f == 3) && w == kLength) {
      if (!get_bytes(k, &sub)) return true;
      Point p;
      ok = ok && dec(sub, &p);
      (f == 2 ? v->nodes : v->nodes0).push_back(p);
      return true;
    }
    return false;
  });
  return r && ok;
}
bool dec(Cursor c, MorphingVelocity* m) {
  bool ok = true;
  const bool r = fields(c, [&](int f, int w, Cursor* k) {
    Cursor sub{};
    if (w != kLength || f < 1 || f > 3) return false;
    if (!get_bytes(k, &sub)) return true;
    Point p;
    ok = ok && dec(sub, &p);
    if (f == 1) {
      m->center = p;
      m->has_center = true;
    } else if (f == 2) {
      m->nodes.push_back(p);
    } else {
      m->freestream = p;
      m->has_freestream = true;
    }
    return true;
  });
  return r && ok;
}
bool dec(Cursor c, ContainerShape* s) {
  bool ok = true;
  const bool r = fields(c, [&](int f, int w, Cursor* k) {
    uint64_t v;
    Cursor sub{};
    if ((f == 1 || f == 2 || f == 3) && w == kVarint) {
      if (!get_varint(k, &v)) return true;
      if (f == 1) s->rows = static_cast<uint32_t>(v);
      if (f == 2) s->cols = static_cast<uint32_t>(v);
      if (f == 3) {
        s->id = static_cast<uint32_t>(v);
        s->has_id = true;
      }
      return true;
    }
    if (f == 4 && w == kLength) {
      if (!get_bytes(k, &sub)) return true;
      ok = ok && dec(sub, &s->origin);
      s->has_origin = true;
      return true;
    }
    if ((f == 5 || f == 6) && w == kFixed64) {
      if (f == 5) s->has_chord = get_double(k, &s->chord);
      if (f == 6) s->has_span = get_double(k, &s->span);
      return true;
    }
    return false;
  });
  return r && ok;
}
bool dec(Cursor c, VortexSheet* s) {
  bool ok = true;
  const bool r = fields(c, [&](int f, int w, Cursor* k) {
    uint64_t v;
    Cursor sub{};
    if (f == 1 && w == kFixed64) {
      s->has_t = get_double(k, &s->t);
      return true;
    }
    if (f >= 2 && f <= 4 && w == kVarint) {
      if (!get_varint(k, &v)) return true;
      (f == 2 ? s->num : f == 3 ? s->rows : s->cols) = static_cast<uint32_t>(v);
      return true;
    }
    if (f == 5 && w == kLength) {
      if (!get_bytes(k, &sub)) return true;
      s->vortices.emplace_back();
      ok = ok && dec(sub, &s->vortices.back());
      return true;
    }
    return false;
  });
  return r && ok;
}

Cursor cursor(const std::string& b) {
  const auto* p = reinterpret_cast<const unsigned char*>(b.data());
  return Cursor{p, p + b.size(), true};
}

}  // namespace

std::string Serialize(const Snapshot2& m) {
  std::string s;
  if (m.has_t) put_double(&s, 1, m.t);
  for (const auto& c : m.container_shapes) put_bytes(&s, 2, enc(c));
  for (const auto& v : m.vortices) put_bytes(&s, 3, enc(v));
  for (const auto& mv : m.morphing_velocities) put_bytes(&s, 4, enc(mv));
  return s;
}

std::string Serialize(const VortexSheet& m) {
  std::string s;
  if (m.has_t) put_double(&s, 1, m.t);
  put_uint(&s, 2, m.num);
  put_uint(&s, 3, m.rows);
  put_uint(&s, 4, m.cols);
  for (const auto& v : m.vortices) put_bytes(&s, 5, enc(v));
  return s;
}

std::string Serialize(const AllVortexSheets& m) {
  std::string s;
  if (m.has_t) put_double(&s, 1, m.t);
  if (m.has_wing) put_bytes(&s, 2, Serialize(m.wing));
  if (m.has_wake) put_bytes(&s, 3, Serialize(m.wake));
  return s;
}

bool Parse(const std::string& bytes, Snapshot2* m) {
  *m = Snapshot2();
  bool ok = true;
  const bool r = fields(cursor(bytes), [&](int f, int w, Cursor* k) {
    Cursor sub{};
    if (f == 1 && w == kFixed64) {
      m->has_t = get_double(k, &m->t);
      return true;
    }
    if (f >= 2 && f <= 4 && w == kLength) {
      if (!get_bytes(k, &sub)) return true;
      if (f == 2) {
        m->container_shapes.emplace_back();
        ok = ok && dec(sub, &m->container_shapes.back());
      } else if (f == 3) {
        m->vortices.emplace_back();
        ok = ok && dec(sub, &m->vortices.back());
      } else {
        m->morphing_velocities.emplace_back();
        ok = ok && dec(sub, &m->morphing_velocities.back());
      }
      return true;
    }
    return false;
  });
  return r && ok;
}

bool Parse(const std::string& bytes, VortexSheet* m) {
  *m = VortexSheet();
  return dec(cursor(bytes), m);
}

bool Parse(const std::string& bytes, AllVortexSheets* m) {
  *m = AllVortexSheets();
  bool ok = true;
  const bool r = fields(cursor(bytes), [&](int f, int w, Cursor* k) {
    Cursor sub{};
    if (f == 1 && w == kFixed64) {
      m->has_t = get_double(k, &m->t);
      return true;
    }
    if ((f == 2 || f == 3) && w == kLength) {
      if (!get_bytes(k, &sub)) return true;
      VortexSheet* s = f == 2 ? &m->wing : &m->wake;
      *s = VortexSheet();
      ok = ok && dec(sub, s);
      (f == 2 ? m->has_wing : m->has_wake) = true;
      return true;
    }
    return false;
  });
  return r && ok;
}

}  // namespace proto

bool RecordWriter::Write(const std::string& message) {
  std::string head;
  uint64_t n = message.size();
  while (n >= 0x80) {
    head.push_back(static_cast<char>((n & 0x7f) | 0x80));
    n >>= 7;
  }
  head.push_back(static_cast<char>(n));
  os_->write(head.data(), static_cast<std::streamsize>(head.size()));
  os_->write(message.data(), static_cast<std::streamsize>(message.size()));
  os_->flush();
  return static_cast<bool>(*os_);
}

bool RecordReader::Read(std::string* message) {
  uint64_t n = 0;
  for (int shift = 0;; shift += 7) {
    const int b = is_->get();
    if (b == std::char_traits<char>::eof()) return false;
    if (shift >= 64) return false;
    n |= static_cast<uint64_t>(b & 0x7f) << shift;
    if (!(b & 0x80)) break;
  }
  message->resize(n);
  if (n) is_->read(&(*message)[0], static_cast<std::streamsize>(n));
  return static_cast<uint64_t>(is_->gcount()) == n || n == 0;
}

}  // namespace UVLM
