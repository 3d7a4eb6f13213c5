// proto_io.h -- the messages of the reference's proto/uvlm.proto that its
// outputs and tools use, as plain structs with their protobuf wire encoding
// written here (protoc and libprotobuf are not part of this build), and the
// record framing of the snapshot streams.
//
// Field numbers and types follow proto/uvlm.proto:4-100.  Serialisation
// writes the fields in field-number order and an optional field only when
// it is set (has_*), as protobuf's C++ serialiser does for proto2; repeated
// messages are one length-delimited field each.  Parsing accepts any field
// order and skips unknown fields.
//
// Record framing: the reference writes its streams with recordio's
// RecordWriter (submodule src/recordio, absent from the reference tree, so
// its byte layout is unknown).  Here a record is the message's byte length
// as a base-128 varint followed by the message (protobuf's delimited form).
#pragma once

#include <cstdint>
#include <istream>
#include <ostream>
#include <string>
#include <vector>

#include "vector3d.h"
#include "wing.h"

namespace UVLM {
namespace proto {

// message VortexRing (uvlm.proto:35-41)
struct VortexRing {
  bool has_gamma = false;
  double gamma = 0.0;
  std::vector<Point> nodes, nodes0;
};

// message MorphingVelocity (uvlm.proto:56-62)
struct MorphingVelocity {
  bool has_center = false, has_freestream = false;
  Point center, freestream;
  std::vector<Point> nodes;
};

// message Snapshot2 (uvlm.proto:66-81)
struct ContainerShape {
  uint32_t rows = 0, cols = 0;
  bool has_id = false, has_origin = false, has_chord = false, has_span = false;
  uint32_t id = 0;
  Point origin;
  double chord = 0.0, span = 0.0;
};
struct Snapshot2 {
  bool has_t = false;
  double t = 0.0;
  std::vector<ContainerShape> container_shapes;
  std::vector<VortexRing> vortices;
  std::vector<MorphingVelocity> morphing_velocities;
};

// message VortexSheet (uvlm.proto:84-93)
struct VortexSheet {
  bool has_t = false;
  double t = 0.0;
  uint32_t num = 0, rows = 0, cols = 0;
  std::vector<VortexRing> vortices;
};

// message AllVortexSheets (uvlm.proto:96-100)
struct AllVortexSheets {
  bool has_t = false, has_wing = false, has_wake = false;
  double t = 0.0;
  VortexSheet wing, wake;
};

std::string Serialize(const Snapshot2& m);
std::string Serialize(const AllVortexSheets& m);
std::string Serialize(const VortexSheet& m);
// false on malformed input (truncated field, bad wire type).
bool Parse(const std::string& bytes, Snapshot2* m);
bool Parse(const std::string& bytes, AllVortexSheets* m);
bool Parse(const std::string& bytes, VortexSheet* m);

}  // namespace proto

// Record streams of serialised messages (the framing above).
class RecordWriter {
 public:
  explicit RecordWriter(std::ostream* os) : os_(os) {}
  bool Write(const std::string& message);
  template <class M>
  bool WriteProtocolMessage(const M& m) {
    return Write(proto::Serialize(m));
  }

 private:
  std::ostream* os_;
};

class RecordReader {
 public:
  explicit RecordReader(std::istream* is) : is_(is) {}
  // false at the end of the stream or on a truncated record.
  bool Read(std::string* message);
  template <class M>
  bool ReadProtocolMessage(M* m) {
    std::string b;
    return Read(&b) && proto::Parse(b, m);
  }

 private:
  std::istream* is_;
};

}  // namespace UVLM
