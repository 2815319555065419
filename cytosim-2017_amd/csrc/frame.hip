// Trajectory frames: Simul::writeObjects (simul_file.cc:402-489) in binary
// mode, with the fiber and solid sections packed on the device from the
// points of the last solve.
//
// The packed sections are laid out in registration order, which is matIndex
// order (fibers, then Solids), so every point has a fixed byte offset: its
// record's offset, plus the record's head (section line for the first record
// of its kind, newline, Object::writeReference, the fiber fields, the point
// count), plus its rank in the record times DIM reals (and a radius for a
// Solid point).  The heads depend only on the registration and the records
// the host gives, so they are built once per layout; a frame is then one
// launch (a thread per head, a thread per point) and one copy to the host.
// CYMG_FRAME_FLOAT converts each real to float as OutputWrapper::writeFloat
// does (iowrapper.cc:535-547); CYMG_FRAME_DOUBLE keeps all 8 bytes.
//
// cymg_read_frame, a reader of such frames, is host code at the end.
#include <cstring>
#include <string>
#include <vector>

#include "context.h"

namespace cymg {

namespace {

template <int RB>
__device__ __forceinline__ void put_real(unsigned char* o, double v)
{
    if constexpr (RB == 8) {
        __builtin_memcpy(o, &v, 8);
    } else {
        const float f = (float)v;   // round to nearest even, as the host conversion
        __builtin_memcpy(o, &f, 4);
    }
}

template <int D, int RB>
__global__ __launch_bounds__(256) void k_frame_pack(const FrameHead* __restrict__ heads, uint32_t nrec,
                                                    const uint32_t* __restrict__ pmap, uint32_t npts,
                                                    const double* __restrict__ X,
                                                    const double* __restrict__ sol_rad, uint32_t sol_first,
                                                    unsigned char* __restrict__ out)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nrec) {
        const FrameHead& h = heads[t];
        unsigned char* o = out + h.off;
        for (uint32_t b = 0; b < h.hlen; ++b) o[b] = h.head[b];
        return;
    }
    t -= nrec;
    if (t >= npts) return;
    const FrameHead& h = heads[pmap[t]];
    const uint32_t stride = (D + h.solid) * RB;
    unsigned char* o = out + h.off + h.hlen + (uint64_t)(t - h.first) * stride;
#pragma unroll
    for (int d = 0; d < D; ++d) put_real<RB>(o + d * RB, X[(size_t)D * t + d]);
    if (h.solid) put_real<RB>(o + D * RB, sol_rad[t - sol_first]);
}

// the bytes of OutputWrapper's binary writes (iowrapper.cc:452-533): native order
struct Bytes {
    std::vector<unsigned char> b;
    void raw(const void* p, size_t n) { b.insert(b.end(), (const unsigned char*)p, (const unsigned char*)p + n); }
    void str(const char* s) { raw(s, std::strlen(s)); }
    void u8(uint32_t v) { const uint8_t x = (uint8_t)v; raw(&x, 1); }
    void u16(uint32_t v) { const uint16_t x = (uint16_t)v; raw(&x, 2); }
    void u32(uint32_t v) { raw(&v, 4); }
    void i32(int32_t v) { raw(&v, 4); }
    void real(double v, int rb)
    {
        if (rb == 8) {
            raw(&v, 8);
        } else {
            const float f = (float)v;
            raw(&f, 4);
        }
    }
    // Object::writeReference (object.cc:76-97)
    void reference(char tag, const cymg_frame_record& r)
    {
        if (r.number <= 65535 && r.prop <= 255 && r.mark == 0) {
            raw(&tag, 1);
            u8(r.prop);
            u16(r.number);
        } else {
            str("$");
            raw(&tag, 1);
            u16(r.prop);
            u32(r.number);
            i32(r.mark);
        }
    }
};

// heads and point map of the current registration (fibers, then Solids)
void frame_layout(Context& c, int rb)
{
    const uint32_t F = c.nfib, S = c.nsol, M = F + S;
    std::vector<FrameHead> heads(M);
    std::vector<uint32_t> pmap(c.npts, 0);
    uint64_t off = 0;
    for (uint32_t m = 0; m < M; ++m) {
        const bool solid = m >= F;
        const cymg_frame_record& r = c.h_frame_rec[m];
        const uint32_t np = solid ? c.h_sol_np[m - F] : c.h_np[m];
        const uint32_t first = solid ? c.h_sol_idx[m - F] : c.h_idx[m];
        if (r.number == 0) throw Error(CYMG_ERR_INVALID, "frame record " + std::to_string(m) + ": number 0");
        if (r.prop > 65535) throw Error(CYMG_ERR_INVALID, "frame record " + std::to_string(m) + ": property index beyond 65535");
        if (np > 65535) throw Error(CYMG_ERR_UNSUPPORTED, "a Mecable of more than 65535 points in a frame");
        Bytes h;
        if (m == 0 && F) h.str("\n#section fiber");   // simul_file.cc:443-447
        if (m == F && S) h.str("\n#section solid");   // simul_file.cc:449-453
        h.str("\n");                                   // ObjectSet::write (object_set.cc:384)
        if (!solid) {
            h.reference('f', r);
            h.u32(r.signature);                        // FiberNaked::write (fiber_naked.cc:1195-1202)
            h.real((double)(np - 1) * c.h_cut[m], rb); // length() = nbSegments() * fnCut
            h.real(r.cut_wished, rb);
            h.real(r.abscissa, rb);
            h.u16(np);                                 // PointSet::write (point_set.cc:337-342)
        } else {
            h.reference('d', r);
            h.u16(np);                                 // Solid::write (solid.cc:896-905)
        }
        FrameHead& fh = heads[m];
        std::memset(&fh, 0, sizeof fh);
        fh.off = off;
        fh.first = first;
        fh.np = np;
        fh.hlen = (uint32_t)h.b.size();
        fh.solid = solid ? 1 : 0;
        std::memcpy(fh.head, h.b.data(), h.b.size());
        for (uint32_t p = 0; p < np; ++p) pmap[first + p] = m;
        off += fh.hlen + (uint64_t)np * (c.dim + fh.solid) * rb;
    }
    c.frame_bytes = off;
    c.d_frame_head.resize(M ? M : 1);
    c.d_frame_pmap.resize(c.npts ? c.npts : 1);
    c.d_frame.resize(off ? off : 1);
    if (M) CYMG_HIP(hipMemcpyAsync(c.d_frame_head.get(), heads.data(), M * sizeof(FrameHead), hipMemcpyHostToDevice, c.stream));
    if (c.npts) CYMG_HIP(hipMemcpyAsync(c.d_frame_pmap.get(), pmap.data(), (size_t)c.npts * 4, hipMemcpyHostToDevice, c.stream));
    CYMG_HIP(hipStreamSynchronize(c.stream));
    c.frame_made_layout = c.layout_gen;
    c.frame_made_rec = c.frame_rec_gen;
    c.frame_made_rb = rb;
}

std::string printf_str(const char* fmt, double t, int a, int b)
{
    char line[160];
    std::snprintf(line, sizeof line, fmt, t, a, b);
    return line;
}

}  // namespace

void write_frame(Context& c, const cymg_frame_params& p, void* out, size_t capacity, size_t* size)
{
    const int rb = p.real_bytes;
    if (rb != CYMG_FRAME_FLOAT && rb != CYMG_FRAME_DOUBLE)
        throw Error(CYMG_ERR_INVALID, "real_bytes must be CYMG_FRAME_FLOAT or CYMG_FRAME_DOUBLE");
    if (c.h_frame_rec.size() != c.nmec())
        throw Error(CYMG_ERR_INVALID, "frame records do not match the registration: cymg_set_frame_records after each registration");
    if ((p.nbefore && !p.before) || (p.nafter && !p.after)) throw Error(CYMG_ERR_INVALID, "null section bytes");
    const char* date = p.date ? p.date : "";
    // frame tag, time line, binary signature (simul_file.cc:415-422; FRAME_TAG simul.h:31;
    // OutputWrapper::writeBinarySignature, iowrapper.cc:378-388)
    Bytes pro;
    pro.str("\n\n#Cytosim  ");
    pro.str(date);
    pro.str((rb == 8 ? printf_str("\n#time %.17g, dim %i, format %i, real 8", p.time, c.dim, 41)
                     : printf_str("\n#time %.6f, dim %i, format %i", p.time, c.dim, 41)).c_str());
    pro.str("\n#binary ");
    pro.u16(12592U);
    pro.str("\n");
    Bytes epi;
    epi.str("\n#section end");
    epi.str("\n#end cytosim ");
    epi.str(date);
    epi.str("\n\n");
    if (c.frame_made_layout != c.layout_gen || c.frame_made_rec != c.frame_rec_gen || c.frame_made_rb != rb)
        frame_layout(c, rb);
    const size_t total = pro.b.size() + p.nbefore + c.frame_bytes + p.nafter + epi.b.size();
    *size = total;
    if (!capacity || !out) return;
    if (capacity < total) throw Error(CYMG_ERR_INVALID, "capacity below the frame's length");
    if (!c.prepared) throw Error(CYMG_ERR_INVALID, "no points: cymg_prepare first");
    unsigned char* o = (unsigned char*)out;
    std::memcpy(o, pro.b.data(), pro.b.size());
    o += pro.b.size();
    if (p.nbefore) std::memcpy(o, p.before, p.nbefore);
    o += p.nbefore;
    const uint32_t M = c.nmec();
    if (c.frame_bytes) {
        const uint32_t nt = M + c.npts;
        const dim3 grid((nt + 255) / 256);
        const uint32_t sol_first = c.nsol ? c.h_sol_idx[0] : 0;
        const double* srad = c.nsol ? c.d_sol_rad.get() : nullptr;
        if (c.dim == 3) {
            if (rb == 8)
                hipLaunchKernelGGL((k_frame_pack<3, 8>), grid, dim3(256), 0, c.stream, c.d_frame_head.get(), M,
                                   c.d_frame_pmap.get(), c.npts, c.d_pos.get(), srad, sol_first, c.d_frame.get());
            else
                hipLaunchKernelGGL((k_frame_pack<3, 4>), grid, dim3(256), 0, c.stream, c.d_frame_head.get(), M,
                                   c.d_frame_pmap.get(), c.npts, c.d_pos.get(), srad, sol_first, c.d_frame.get());
        } else {
            if (rb == 8)
                hipLaunchKernelGGL((k_frame_pack<2, 8>), grid, dim3(256), 0, c.stream, c.d_frame_head.get(), M,
                                   c.d_frame_pmap.get(), c.npts, c.d_pos.get(), srad, sol_first, c.d_frame.get());
            else
                hipLaunchKernelGGL((k_frame_pack<2, 4>), grid, dim3(256), 0, c.stream, c.d_frame_head.get(), M,
                                   c.d_frame_pmap.get(), c.npts, c.d_pos.get(), srad, sol_first, c.d_frame.get());
        }
        CYMG_HIP(hipGetLastError());
        CYMG_HIP(hipMemcpyAsync(o, c.d_frame.get(), c.frame_bytes, hipMemcpyDeviceToHost, c.stream));
        CYMG_HIP(hipStreamSynchronize(c.stream));
    }
    o += c.frame_bytes;
    if (p.nafter) std::memcpy(o, p.after, p.nafter);
    o += p.nafter;
    std::memcpy(o, epi.b.data(), epi.b.size());
}

}  // namespace cymg

// ---------------------------------------------------------------------------
// Reader: Simul::readObjects (simul_file.cc:240-390) for the fiber and solid
// sections, with Object::readReference (object.cc:109-152), FiberNaked::read
// (fiber_naked.cc:1208-1240), PointSet::read (point_set.cc:346-366),
// Solid::read (solid.cc:908-925) and InputWrapper's binary reads
// (iowrapper.cc:42-283).  Records of other sections are skipped to the next
// line that opens with '#'.

namespace {

struct ReadError {
    int code;
    std::string msg;
};

struct Reader {
    const unsigned char* d;
    size_t n, p;
    int real = 4, fdim = 3;
    bool binary = false, swap = false;

    int getc() { return p < n ? d[p++] : -1; }
    void take(void* out, size_t k)
    {
        if (p + k > n) throw ReadError{ CYMG_ERR_INVALID, "frame truncated" };
        std::memcpy(out, d + p, k);
        if (swap) {
            unsigned char* b = (unsigned char*)out;
            for (size_t i = 0; i < k / 2; ++i) std::swap(b[i], b[k - 1 - i]);
        }
        p += k;
    }
    std::string line()
    {
        const void* q = p < n ? std::memchr(d + p, '\n', n - p) : nullptr;
        const size_t e = q ? (size_t)((const unsigned char*)q - d) : n;
        std::string s((const char*)d + p, e - p);
        p = q ? e + 1 : n;
        return s;
    }
    uint32_t u8() { uint8_t v; take(&v, 1); return v; }
    uint32_t u16() { uint16_t v; take(&v, 2); return v; }
    uint32_t u32() { uint32_t v; take(&v, 4); return v; }
    int32_t i32() { int32_t v; take(&v, 4); return v; }
    double real_()
    {
        if (real == 8) { double v; take(&v, 8); return v; }
        float f; take(&f, 4); return (double)f;
    }
};

}  // namespace

extern "C" int cymg_read_frame(const void* data, size_t size, int dim, cymg_frame_object* objects,
                               uint32_t max_objects, uint32_t* nobjects, double* points, double* radii,
                               size_t max_points, size_t* npoints, double* time, size_t* used)
{
    try {
        if (dim != 2 && dim != 3) throw ReadError{ CYMG_ERR_INVALID, "dim must be 2 or 3" };
        if (!data || !nobjects || !npoints) throw ReadError{ CYMG_ERR_INVALID, "null argument" };
        Reader in{ (const unsigned char*)data, size, 0 };
        uint32_t no = 0;
        size_t np = 0;
        int c = '\n';
        std::string section;
        bool started = false, ended = false;
        double t = 0;
        while (true) {
            int tag = in.getc();
            while (tag >= 0 && (tag == ' ' || tag == '\n' || tag == '\t' || tag == '\r' || tag == '\v' || tag == '\f')) {
                c = tag;
                tag = in.getc();
            }
            if (tag < 0) break;
            if (c != '\n') {
                in.line();
                c = '\n';
                continue;
            }
            int pretag = 0;
            if (tag == '$') {
                pretag = tag;
                tag = in.getc();
            }
            if (tag == '#') {
                const std::string l = in.line();
                if (l.compare(0, 8, "Cytosim ") == 0) {
                    started = true;
                } else if (l.compare(0, 8, "section ") == 0) {
                    section = l.substr(8);
                } else if (l.compare(0, 5, "time ") == 0) {
                    int f = 41, dd = 3;
                    std::sscanf(l.c_str(), "time %lf, dim %i, format %i", &t, &dd, &f);
                    if (dd != 2 && dd != 3) throw ReadError{ CYMG_ERR_INVALID, "frame of dimension " + std::to_string(dd) };
                    in.fdim = dd;
                    const size_t k = l.rfind(", real 8");
                    in.real = (k != std::string::npos && k + 8 == l.size()) ? 8 : 4;
                } else if (l.compare(0, 7, "binary ") == 0) {
                    // InputWrapper::setBinarySwap (iowrapper.cc:42-48): "01" as written natively
                    in.binary = true;
                    in.swap = l.size() > 7 && l[7] != '0';
                } else if (l.compare(0, 12, "end cytosim ") == 0) {
                    ended = true;
                    break;
                }
                continue;
            }
            if (!in.binary) throw ReadError{ CYMG_ERR_UNSUPPORTED, "only binary frames are read" };
            const bool fib = section == "fiber" && tag == 'f';
            const bool sol = section == "solid" && tag == 'd';
            if (!fib && !sol) {
                // a record of another section: skip to the next line that opens with '#'
                size_t q = in.p;
                while (q + 1 < in.n && !(in.d[q] == '\n' && in.d[q + 1] == '#')) ++q;
                in.p = q + 1 < in.n ? q : in.n;
                c = '\n';
                continue;
            }
            cymg_frame_object o{};
            o.tag = (char)tag;
            if (pretag == '$') {
                o.prop = in.u16();
                o.number = in.u32();
                o.mark = in.i32();
            } else {
                o.prop = in.u8();
                o.number = in.u16();
                o.mark = 0;
            }
            if (o.number == 0) throw ReadError{ CYMG_ERR_INVALID, "Invalid (null) object reference" };
            if (fib) {
                o.signature = in.u32();
                o.length = in.real_();
                o.cut_wished = in.real_();
                o.abscissa = in.real_();
            }
            o.points = in.u16();
            o.first = np;
            for (uint32_t k = 0; k < o.points; ++k, ++np) {
                // InputWrapper::readFloatVector (iowrapper.cc:261-287)
                double v[3] = { 0, 0, 0 };
                for (int d = 0; d < in.fdim; ++d) {
                    const double x = in.real_();
                    if (d < dim) v[d] = x;
                }
                const double r = sol ? in.real_() : 0.0;
                if (points && np < max_points)
                    for (int d = 0; d < dim; ++d) points[(size_t)dim * np + d] = v[d];
                if (radii && np < max_points) radii[np] = r;
            }
            if (objects && no < max_objects) objects[no] = o;
            ++no;
        }
        if (!started) throw ReadError{ CYMG_ERR_INVALID, "no frame" };
        if (!ended) throw ReadError{ CYMG_ERR_INVALID, "frame without its end line" };
        *nobjects = no;
        *npoints = np;
        if (time) *time = t;
        if (used) *used = in.p;
        return CYMG_OK;
    } catch (const ReadError& e) {
        cymg::set_thread_error(e.msg);
        return e.code;
    } catch (const std::exception& e) {
        cymg::set_thread_error(e.what());
        return CYMG_ERR_INVALID;
    }
}

namespace cymg {

static thread_local std::string g_thread_error = "null context";

const char* thread_error() { return g_thread_error.c_str(); }
void set_thread_error(const std::string& msg) { g_thread_error = msg; }

}  // namespace cymg
