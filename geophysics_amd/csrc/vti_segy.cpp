// vti_segy.cpp — shot records to Seismic Unix (SU) and SEG-Y files (host).
//
// vti_bin2su has the semantics of utils-binary2segy/binary2segy.c:43-85:
// every trace of the output is the 240-byte header of the matching trace of
// a template SU file (copied as it is) followed by ns samples read from a
// raw float file, each sample's four bytes reversed (swap_float_4 :136-144).
// Unlike the reference it checks every open, read and write.
//
// vti_su_write / vti_segy_write write records as the library produces them
// (trace-major: ntr traces of ns samples per shot, the layout of the
// drop-ins' shot files) with headers built from the acquisition geometry.
// The 240-byte trace header is the SU/SEG-Y layout (the key offsets below,
// e.g. tracl at byte 0, ns at 114, dt at 116: utils-binary2segy/hc/cjbsegy.h
// declares the same fields in this order); SU files keep the host's byte
// order, SEG-Y files are big-endian throughout, with a 3200-byte EBCDIC
// textual header, a 400-byte binary header and IEEE float samples (format
// code 5, SEG-Y revision 1).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/vti.h"
#include "vti_internal.h"

namespace {

constexpr int kHdr = 240;

// byte offsets of the trace-header keys written here
enum : int {
    K_TRACL = 0, K_TRACR = 4, K_FLDR = 8, K_TRACF = 12, K_EP = 16, K_TRID = 28, K_OFFSET = 36, K_GELEV = 40,
    K_SDEPTH = 48, K_SCALEL = 68, K_SCALCO = 70, K_SX = 72, K_SY = 76, K_GX = 80, K_GY = 84, K_COUNIT = 88,
    K_NS = 114, K_DT = 116
};

struct File {
    FILE *f = nullptr;
    ~File()
    {
        if (f) std::fclose(f);
    }
};

int io_fail(const std::string &m) { return vti::fail(VTI_ERR_IO, m); }

void put32(uint8_t *h, int off, int32_t v, bool big)
{
    const uint32_t u = (uint32_t)v;
    for (int i = 0; i < 4; ++i) h[off + i] = (uint8_t)(big ? u >> (24 - 8 * i) : u >> (8 * i));
}

void put16(uint8_t *h, int off, int v, bool big)
{
    const uint16_t u = (uint16_t)v;
    h[off + (big ? 0 : 1)] = (uint8_t)(u >> 8);
    h[off + (big ? 1 : 0)] = (uint8_t)u;
}

void swap4(void *p, size_t n)
{
    uint8_t *b = static_cast<uint8_t *>(p);
    for (size_t i = 0; i < n; ++i, b += 4) {
        std::swap(b[0], b[3]);
        std::swap(b[1], b[2]);
    }
}

bool host_big()
{
    const uint16_t one = 1;
    uint8_t b;
    std::memcpy(&b, &one, 1);
    return b == 0;
}

// ASCII -> EBCDIC (code page 037) for the characters of the textual header
uint8_t ebcdic(char c)
{
    static const char *a = " .<(+|&!$*);-/,%_>?:#@'=\"abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ0123456789";
    static const uint8_t e[] = {
        0x40, 0x4B, 0x4C, 0x4D, 0x4E, 0x4F, 0x50, 0x5A, 0x5B, 0x5C, 0x5D, 0x5E, 0x60, 0x61, 0x6B, 0x6C, 0x6D, 0x6E,
        0x6F, 0x7A, 0x7B, 0x7C, 0x7D, 0x7E, 0x7F,
        0x81, 0x82, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x91, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99,
        0xA2, 0xA3, 0xA4, 0xA5, 0xA6, 0xA7, 0xA8, 0xA9,
        0xC1, 0xC2, 0xC3, 0xC4, 0xC5, 0xC6, 0xC7, 0xC8, 0xC9, 0xD1, 0xD2, 0xD3, 0xD4, 0xD5, 0xD6, 0xD7, 0xD8, 0xD9,
        0xE2, 0xE3, 0xE4, 0xE5, 0xE6, 0xE7, 0xE8, 0xE9,
        0xF0, 0xF1, 0xF2, 0xF3, 0xF4, 0xF5, 0xF6, 0xF7, 0xF8, 0xF9};
    const char *p = std::strchr(a, c);
    return (c && p) ? e[p - a] : 0x40;
}

bool geometry_ok(const vti_su_geometry *g)
{
    return g && g->nshot >= 1 && g->ntr >= 1 && g->ns >= 1 && g->ns <= 65535 && g->dt_us >= 1 &&
           g->dt_us <= 65535 && g->scalco != 0;
}

// the trace header of trace it of shot is (global trace k)
void trace_header(uint8_t *h, const vti_su_geometry &g, int is, int it, long long k, bool big)
{
    std::memset(h, 0, kHdr);
    const double sx = g.sx0 + (double)is * g.dsx, gx = g.gx0 + (double)is * g.dgx_shot + (double)it * g.dgx;
    const double sy = g.sy0 + (double)is * g.dsy, gy = g.gy0 + (double)is * g.dgy_shot + (double)it * g.dgy;
    // coordinates are stored as integers; scalco < 0 divides by -scalco, > 0 multiplies
    auto co = [&](double v) {
        const double s = g.scalco < 0 ? v * -g.scalco : v / g.scalco;
        return (int32_t)(s < 0 ? s - 0.5 : s + 0.5);
    };
    put32(h, K_TRACL, (int32_t)(k + 1), big);
    put32(h, K_TRACR, (int32_t)(k + 1), big);
    put32(h, K_FLDR, g.fldr0 + is, big);
    put32(h, K_TRACF, it + 1, big);
    put32(h, K_EP, g.fldr0 + is, big);
    put16(h, K_TRID, 1, big);
    const double dxo = gx - sx, dyo = gy - sy;
    const double off = dyo == 0.0 ? dxo : (dxo < 0 ? -1.0 : 1.0) * std::sqrt(dxo * dxo + dyo * dyo);
    put32(h, K_OFFSET, (int32_t)(off < 0 ? off - 0.5 : off + 0.5), big);
    put32(h, K_GELEV, (int32_t)-(g.gz >= 0 ? (int32_t)(g.gz + 0.5) : (int32_t)(g.gz - 0.5)), big);
    put32(h, K_SDEPTH, (int32_t)(g.sz >= 0 ? g.sz + 0.5 : g.sz - 0.5), big);
    put16(h, K_SCALEL, 1, big);
    put16(h, K_SCALCO, g.scalco, big);
    put32(h, K_SX, co(sx), big);
    put32(h, K_SY, co(sy), big);
    put32(h, K_GX, co(gx), big);
    put32(h, K_GY, co(gy), big);
    put16(h, K_COUNIT, 1, big);
    put16(h, K_NS, g.ns, big);
    put16(h, K_DT, g.dt_us, big);
}

int write_traces(FILE *f, const float *rec, const vti_su_geometry &g, bool big, const char *who)
{
    std::vector<uint8_t> h(kHdr);
    std::vector<float> tr(g.ns);
    const bool swap = big != host_big();
    long long k = 0;
    for (int is = 0; is < g.nshot; ++is)
        for (int it = 0; it < g.ntr; ++it, ++k) {
            trace_header(h.data(), g, is, it, k, big);
            std::memcpy(tr.data(), rec + (size_t)k * g.ns, (size_t)g.ns * 4);
            if (swap) swap4(tr.data(), tr.size());
            if (std::fwrite(h.data(), 1, kHdr, f) != (size_t)kHdr ||
                std::fwrite(tr.data(), 4, tr.size(), f) != tr.size())
                return io_fail(std::string(who) + ": write failed");
        }
    return VTI_OK;
}

}  // namespace

extern "C" {

int vti_bin2su(const char *su_template, const char *raw, const char *out, long long ntrace, int ns, int swap)
{
    if (!su_template || !raw || !out || ntrace < 0 || ns < 1)
        return vti::fail(VTI_ERR_ARG, "vti_bin2su: bad arguments");
    File t, r, o;
    if (!(t.f = std::fopen(su_template, "rb"))) return io_fail(std::string("vti_bin2su: cannot open ") + su_template);
    if (!(r.f = std::fopen(raw, "rb"))) return io_fail(std::string("vti_bin2su: cannot open ") + raw);
    if (!(o.f = std::fopen(out, "wb"))) return io_fail(std::string("vti_bin2su: cannot open ") + out);
    std::vector<uint8_t> h(kHdr);
    std::vector<float> tr(ns);
    for (long long k = 0; k < ntrace; ++k) {
        // the template trace: its header is kept, its samples are skipped
        if (std::fread(h.data(), 1, kHdr, t.f) != (size_t)kHdr || std::fseek(t.f, (long)ns * 4, SEEK_CUR) != 0)
            return io_fail("vti_bin2su: template ends at trace " + std::to_string(k));
        if (std::fread(tr.data(), 4, tr.size(), r.f) != tr.size())
            return io_fail("vti_bin2su: raw data ends at trace " + std::to_string(k));
        if (swap) swap4(tr.data(), tr.size());
        if (std::fwrite(h.data(), 1, kHdr, o.f) != (size_t)kHdr || std::fwrite(tr.data(), 4, tr.size(), o.f) != tr.size())
            return io_fail("vti_bin2su: write failed");
    }
    return VTI_OK;
}

int vti_su_write(const char *path, const float *rec, const vti_su_geometry *g, int big_endian)
{
    if (!path || !rec || !geometry_ok(g)) return vti::fail(VTI_ERR_ARG, "vti_su_write: bad arguments");
    File f;
    if (!(f.f = std::fopen(path, "wb"))) return io_fail(std::string("vti_su_write: cannot open ") + path);
    return write_traces(f.f, rec, *g, big_endian != 0, "vti_su_write");
}

int vti_segy_write(const char *path, const float *rec, const vti_su_geometry *g, const char *text)
{
    if (!path || !rec || !geometry_ok(g) || g->ntr > 32767)
        return vti::fail(VTI_ERR_ARG, "vti_segy_write: bad arguments");
    File f;
    if (!(f.f = std::fopen(path, "wb"))) return io_fail(std::string("vti_segy_write: cannot open ") + path);
    // textual header: 40 card images of 80 characters, "C nn " then the
    // caller's text (lines split at '\n'), EBCDIC
    std::vector<uint8_t> txt(3200, 0x40);
    const char *p = text ? text : "";
    for (int line = 0; line < 40; ++line) {
        char card[81];
        std::snprintf(card, sizeof(card), "C%2d ", line + 1);
        int c = 4;
        while (*p && *p != '\n' && c < 80) card[c++] = *p++;
        while (*p && *p != '\n') ++p;   // the rest of an over-long line is dropped
        if (*p == '\n') ++p;
        for (int i = 0; i < c; ++i) txt[line * 80 + i] = ebcdic(card[i]);
    }
    // binary header (big-endian): job, line and reel numbers 1, traces per
    // record, sample interval and count, format 5 (IEEE float), sorting 1
    // (as recorded), metres, SEG-Y revision 1.0, fixed trace length
    std::vector<uint8_t> bh(400, 0);
    put32(bh.data(), 0, 1, true);
    put32(bh.data(), 4, 1, true);
    put32(bh.data(), 8, 1, true);
    put16(bh.data(), 12, g->ntr, true);
    put16(bh.data(), 16, g->dt_us, true);
    put16(bh.data(), 18, g->dt_us, true);
    put16(bh.data(), 20, g->ns, true);
    put16(bh.data(), 22, g->ns, true);
    put16(bh.data(), 24, 5, true);
    put16(bh.data(), 28, 1, true);
    put16(bh.data(), 54, 1, true);
    put16(bh.data(), 300, 0x0100, true);
    put16(bh.data(), 302, 1, true);
    if (std::fwrite(txt.data(), 1, txt.size(), f.f) != txt.size() || std::fwrite(bh.data(), 1, bh.size(), f.f) != bh.size())
        return io_fail("vti_segy_write: write failed");
    return write_traces(f.f, rec, *g, true, "vti_segy_write");
}

}  // extern "C"
