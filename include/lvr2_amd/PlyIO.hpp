// lvr2::hip_ply: PLY point-cloud ingestion and mesh output for the HIP path, in place of what
// ModelFactory::readModel / saveModel do through PLYIO for .ply files (io/ModelFactory.cpp,
// io/modelio/PLYIO.cpp).
//   - readPly: the "vertex" element (or "point", which PLYIO reads when a file has no vertex
//     element with x, y, z, PLYIO.cpp:680-715): x, y, z as the points; nx, ny, nz as the normals
//     when all three are present; "face" vertex_indices as triangles.  Other properties are
//     skipped.  Formats: ascii, binary_little_endian, binary_big_endian; every PLY scalar type
//     (char/int8 ... double/float64) for any property, list properties of any item type (at most
//     10^6 items a row).  Counts the file is too short for are refused before any allocation.
//     Values are converted to float; indices to int32.
//   - writePly: binary_little_endian, "vertex" with float x, y, z (and nx, ny, nz when given),
//     "face" with a uchar-counted int list vertex_indices, as PLYIO::save writes a mesh
//     (PLYIO.cpp:160-250).
// Errors throw std::runtime_error naming the file and the problem.
#ifndef LVR2_AMD_PLYIO_HPP
#define LVR2_AMD_PLYIO_HPP

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace lvr2
{
namespace hip_ply
{

struct Cloud {
    std::vector<float> points;   // xyz per point
    std::vector<float> normals;  // xyz per point, or empty
    std::vector<int32_t> faces;  // 3 vertex ids per triangle, or empty
    size_t size() const { return points.size() / 3; }
};

namespace detail
{

enum class Fmt { Ascii, Le, Be };

struct Prop {
    std::string name;
    int type = 0;        // bytes of a scalar, sign and kind in kind
    char kind = 'f';     // 'i' signed, 'u' unsigned, 'f' float
    bool list = false;
    int ctype = 0;       // count type (list)
    char ckind = 'u';
};

struct Element {
    std::string name;
    size_t count = 0;
    std::vector<Prop> props;
};

inline bool scalar_type(const std::string& t, int& bytes, char& kind)
{
    static const struct {
        const char* n;
        int b;
        char k;
    } T[] = {{"char", 1, 'i'},   {"int8", 1, 'i'},    {"uchar", 1, 'u'},   {"uint8", 1, 'u'},  {"short", 2, 'i'},
             {"int16", 2, 'i'},  {"ushort", 2, 'u'},  {"uint16", 2, 'u'},  {"int", 4, 'i'},    {"int32", 4, 'i'},
             {"uint", 4, 'u'},   {"uint32", 4, 'u'},  {"float", 4, 'f'},   {"float32", 4, 'f'}, {"double", 8, 'f'},
             {"float64", 8, 'f'}};
    for (const auto& e : T)
        if (t == e.n) {
            bytes = e.b;
            kind = e.k;
            return true;
        }
    return false;
}

inline double decode(const unsigned char* p, int bytes, char kind, bool swap)
{
    unsigned char b[8];
    for (int i = 0; i < bytes; i++) b[i] = swap ? p[bytes - 1 - i] : p[i];
    switch (kind) {
    case 'f':
        if (bytes == 4) {
            float f;
            std::memcpy(&f, b, 4);
            return f;
        } else {
            double d;
            std::memcpy(&d, b, 8);
            return d;
        }
    case 'i':
        if (bytes == 1) return (double)(int8_t)b[0];
        if (bytes == 2) {
            int16_t v;
            std::memcpy(&v, b, 2);
            return v;
        } else {
            int32_t v;
            std::memcpy(&v, b, 4);
            return v;
        }
    default:
        if (bytes == 1) return (double)b[0];
        if (bytes == 2) {
            uint16_t v;
            std::memcpy(&v, b, 2);
            return v;
        } else {
            uint32_t v;
            std::memcpy(&v, b, 4);
            return v;
        }
    }
}

inline bool host_big_endian()
{
    const uint16_t one = 1;
    unsigned char c;
    std::memcpy(&c, &one, 1);
    return c == 0;
}

}  // namespace detail

inline Cloud readPly(const std::string& path)
{
    using namespace detail;
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("readPly: cannot open " + path);
    std::string line;
    std::getline(in, line);
    if (line.substr(0, 3) != "ply") throw std::runtime_error("readPly: " + path + " is not a PLY file");
    Fmt fmt = Fmt::Ascii;
    std::vector<Element> els;
    bool have_format = false;
    for (;;) {
        if (!std::getline(in, line)) throw std::runtime_error("readPly: " + path + ": header without end_header");
        if (!line.empty() && line.back() == '\r') line.pop_back();
        std::istringstream ls(line);
        std::string w;
        ls >> w;
        if (w == "format") {
            std::string f;
            ls >> f;
            if (f == "ascii")
                fmt = Fmt::Ascii;
            else if (f == "binary_little_endian")
                fmt = Fmt::Le;
            else if (f == "binary_big_endian")
                fmt = Fmt::Be;
            else
                throw std::runtime_error("readPly: " + path + ": unknown format " + f);
            have_format = true;
        } else if (w == "element") {
            Element e;
            long long c = -1;
            ls >> e.name >> c;
            if (c < 0) throw std::runtime_error("readPly: " + path + ": bad element line: " + line);
            e.count = (size_t)c;
            els.push_back(e);
        } else if (w == "property") {
            if (els.empty()) throw std::runtime_error("readPly: " + path + ": property before any element");
            Prop p;
            std::string t;
            ls >> t;
            if (t == "list") {
                std::string ct, it;
                ls >> ct >> it >> p.name;
                p.list = true;
                if (!scalar_type(ct, p.ctype, p.ckind) || !scalar_type(it, p.type, p.kind) || p.ckind == 'f')
                    throw std::runtime_error("readPly: " + path + ": bad list property: " + line);
            } else {
                ls >> p.name;
                if (!scalar_type(t, p.type, p.kind))
                    throw std::runtime_error("readPly: " + path + ": unknown property type " + t);
            }
            els.back().props.push_back(p);
        } else if (w == "end_header") {
            break;
        }
        // comment, obj_info: skipped
    }
    if (!have_format) throw std::runtime_error("readPly: " + path + ": no format line");
    const bool swap = fmt != Fmt::Ascii && ((fmt == Fmt::Be) != host_big_endian());
    // which element holds the points: "vertex" with x, y, z, else "point"
    auto has_xyz = [](const Element& e) {
        int c = 0;
        for (const auto& p : e.props) c += (!p.list && (p.name == "x" || p.name == "y" || p.name == "z")) ? 1 : 0;
        return c == 3;
    };
    int src = -1;
    for (size_t i = 0; i < els.size(); i++)
        if (els[i].name == "vertex" && has_xyz(els[i])) src = (int)i;
    if (src < 0)
        for (size_t i = 0; i < els.size(); i++)
            if (els[i].name == "point" && has_xyz(els[i])) src = (int)i;
    if (src < 0) throw std::runtime_error("readPly: " + path + ": no vertex or point element with x, y, z");

    // counts the file cannot hold are refused before anything is allocated: a row needs at
    // least its scalar bytes (binary) or two characters per value (ascii)
    {
        const std::streampos here = in.tellg();
        in.seekg(0, std::ios::end);
        const double left = (double)(in.tellg() - here);
        in.seekg(here);
        double need = 0.0;
        for (const auto& e : els) {
            double row = 0.0;
            for (const auto& p : e.props) row += fmt == Fmt::Ascii ? 2.0 : (double)(p.list ? p.ctype : p.type);
            need += row * (double)e.count;
        }
        if (fmt == Fmt::Ascii && need > 0.0) need -= 1.0;  // the last value needs no separator
        if (need > left) throw std::runtime_error("readPly: " + path + ": truncated data (the header promises more)");
    }

    Cloud out;
    std::vector<unsigned char> buf;
    for (size_t ei = 0; ei < els.size(); ei++) {
        const Element& e = els[ei];
        int slot[6] = {-1, -1, -1, -1, -1, -1};  // x y z nx ny nz
        int face_prop = -1;
        for (size_t k = 0; k < e.props.size(); k++) {
            const auto& p = e.props[k];
            static const char* names[6] = {"x", "y", "z", "nx", "ny", "nz"};
            for (int j = 0; j < 6; j++)
                if (!p.list && p.name == names[j]) slot[j] = (int)k;
            if (p.list && (p.name == "vertex_indices" || p.name == "vertex_index")) face_prop = (int)k;
        }
        const bool points = (int)ei == src;
        const bool normals = points && slot[3] >= 0 && slot[4] >= 0 && slot[5] >= 0;
        const bool faces = e.name == "face" && face_prop >= 0;
        if (points) {
            out.points.resize(3 * e.count);
            if (normals) out.normals.resize(3 * e.count);
        }
        std::vector<double> vals(e.props.size());
        std::vector<double> list;
        for (size_t r = 0; r < e.count; r++) {
            for (size_t k = 0; k < e.props.size(); k++) {
                const auto& p = e.props[k];
                if (fmt == Fmt::Ascii) {
                    if (p.list) {
                        double cnt;
                        if (!(in >> cnt) || cnt < 0 || cnt > 1e6)
                            throw std::runtime_error("readPly: " + path + ": truncated list");
                        list.resize((size_t)cnt);
                        for (auto& v : list)
                            if (!(in >> v)) throw std::runtime_error("readPly: " + path + ": truncated list");
                    } else if (!(in >> vals[k])) {
                        throw std::runtime_error("readPly: " + path + ": truncated data");
                    }
                } else {
                    if (p.list) {
                        buf.resize(p.ctype);
                        if (!in.read((char*)buf.data(), p.ctype))
                            throw std::runtime_error("readPly: " + path + ": truncated list");
                        const double dc = decode(buf.data(), p.ctype, p.ckind, swap);
                        if (dc < 0 || dc > 1e6) throw std::runtime_error("readPly: " + path + ": bad list count");
                        const size_t cnt = (size_t)dc;
                        buf.resize(cnt * p.type);
                        if (cnt && !in.read((char*)buf.data(), (std::streamsize)buf.size()))
                            throw std::runtime_error("readPly: " + path + ": truncated list");
                        list.resize(cnt);
                        for (size_t j = 0; j < cnt; j++) list[j] = decode(buf.data() + j * p.type, p.type, p.kind, swap);
                    } else {
                        unsigned char b[8];
                        if (!in.read((char*)b, p.type)) throw std::runtime_error("readPly: " + path + ": truncated data");
                        vals[k] = decode(b, p.type, p.kind, swap);
                    }
                }
                if (faces && (int)k == face_prop) {
                    // a polygon is split into a fan of triangles
                    for (size_t j = 2; j < list.size(); j++) {
                        out.faces.push_back((int32_t)list[0]);
                        out.faces.push_back((int32_t)list[j - 1]);
                        out.faces.push_back((int32_t)list[j]);
                    }
                }
            }
            if (points) {
                for (int j = 0; j < 3; j++) out.points[3 * r + j] = (float)vals[slot[j]];
                if (normals)
                    for (int j = 0; j < 3; j++) out.normals[3 * r + j] = (float)vals[slot[3 + j]];
            }
        }
    }
    return out;
}

inline void writePly(const std::string& path, const float* points, size_t n, const float* normals,
                     const int32_t* faces, size_t nfaces)
{
    std::ofstream o(path, std::ios::binary);
    if (!o) throw std::runtime_error("writePly: cannot create " + path);
    o << "ply\nformat binary_little_endian 1.0\nelement vertex " << n << "\n"
      << "property float x\nproperty float y\nproperty float z\n";
    if (normals) o << "property float nx\nproperty float ny\nproperty float nz\n";
    if (faces && nfaces) o << "element face " << nfaces << "\nproperty list uchar int vertex_indices\n";
    o << "end_header\n";
    const bool swap = detail::host_big_endian();
    auto put32 = [&](const void* v) {
        unsigned char b[4];
        std::memcpy(b, v, 4);
        if (swap) std::swap(b[0], b[3]), std::swap(b[1], b[2]);
        o.write((const char*)b, 4);
    };
    for (size_t i = 0; i < n; i++) {
        for (int j = 0; j < 3; j++) put32(points + 3 * i + j);
        if (normals)
            for (int j = 0; j < 3; j++) put32(normals + 3 * i + j);
    }
    if (faces)
        for (size_t f = 0; f < nfaces; f++) {
            const unsigned char three = 3;
            o.write((const char*)&three, 1);
            for (int j = 0; j < 3; j++) put32(faces + 3 * f + j);
        }
    if (!o) throw std::runtime_error("writePly: write failed on " + path);
}

}  // namespace hip_ply
}  // namespace lvr2

#endif
