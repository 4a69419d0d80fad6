// Node-set files without netCDF: a reader for the HDF5 layout of the reference's
// nodesets/*.nc (written by netCDF-4; read by get_nodeset_from_file, nodesets.c:18-90).
//
// A node file holds a group "nodeset" with five 1-D float64 datasets x, y, z, lambda, phi of
// the same length N (the root dimension "hid").  The reader follows exactly what those files
// use and refuses anything else with a message, never by guessing:
//   - superblock version 0 or 1 (root group entry) or 2 or 3 (root object header address),
//     8-byte offsets and lengths;
//   - version-2 object headers ("OHDR", continuation blocks "OCHK");
//   - groups whose links are stored in the object header (link messages), hard links only;
//   - datasets with a version 1 or 2 dataspace of rank 1, a little-endian IEEE float64
//     datatype, and contiguous storage (layout message version 3 or 4).
// Compact, chunked or filtered datasets, version-1 object headers, symbol-table groups and
// dense link storage are reported as unsupported.  Every read is bounds-checked against the file
// size, so a truncated or damaged file gives an error, not a wild read.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/serbf.h"

namespace serbf {
void set_error(const std::string &msg);   // the message serbf_last_error returns
}

namespace {

struct H5File {
    std::vector<unsigned char> b;
    uint64_t base = 0;
    std::string err;

    bool fail(const std::string &m)
    {
        if (err.empty()) err = m;
        return false;
    }
    bool have(uint64_t off, uint64_t len) const { return off <= b.size() && len <= b.size() - off; }
    uint64_t u(uint64_t off, int nbytes) const
    {
        uint64_t v = 0;
        for (int i = nbytes - 1; i >= 0; i--) v = (v << 8) | b[off + i];
        return v;
    }
};

struct Message {
    int type;
    uint64_t off, size;   // body within the file
};

constexpr uint64_t kUndef = ~0ull;

// messages of a version-2 object header, continuation blocks included
bool read_header(H5File &f, uint64_t addr, std::vector<Message> &out)
{
    out.clear();
    if (addr == kUndef || addr > ~0ull - f.base) return f.fail("object header address undefined");
    addr += f.base;
    if (!f.have(addr, 6)) return f.fail("object header outside the file");
    if (std::memcmp(&f.b[addr], "OHDR", 4) != 0) {
        if (f.b[addr] == 1) return f.fail("version-1 object headers are not supported");
        return f.fail("object header signature missing");
    }
    if (f.b[addr + 4] != 2) return f.fail("object header version is not 2");
    const unsigned flags = f.b[addr + 5];
    uint64_t p = addr + 6;
    if (flags & 0x20) p += 16;   // access, modification, change, birth times
    if (flags & 0x10) p += 4;    // attribute phase-change values
    const int szb = 1 << (flags & 3);
    if (!f.have(p, szb)) return f.fail("object header truncated");
    const uint64_t csz = f.u(p, szb);
    p += szb;
    const int hdr = (flags & 0x04) ? 6 : 4;   // type, size, flags [, creation order]
    struct Block {
        uint64_t s, e;
    };
    std::vector<Block> blocks;
    if (!f.have(p, csz)) return f.fail("object header chunk outside the file");
    blocks.push_back({p, p + csz});
    for (size_t bi = 0; bi < blocks.size(); bi++) {
        if (blocks.size() > 64) return f.fail("too many object header continuation blocks");
        uint64_t q = blocks[bi].s;
        const uint64_t e = blocks[bi].e;
        while (e - q >= (uint64_t)hdr) {
            const int type = f.b[q];
            const uint64_t sz = f.u(q + 1, 2);
            q += hdr;
            if (sz > e - q) return f.fail("object header message overruns its chunk");
            out.push_back({type, q, sz});
            if (type == 0x10) {   // continuation: offset, length of the next block
                if (sz < 16) return f.fail("continuation message too short");
                const uint64_t ca = f.u(q, 8), cl = f.u(q + 8, 8);
                if (ca == kUndef || ca > ~0ull - f.base) return f.fail("continuation address undefined");
                const uint64_t a = ca + f.base;
                if (cl < 8 || !f.have(a, cl)) return f.fail("continuation block outside the file");
                if (std::memcmp(&f.b[a], "OCHK", 4) != 0) return f.fail("continuation block signature missing");
                blocks.push_back({a + 4, a + cl - 4});   // between the signature and the checksum
            }
            q += sz;
        }
    }
    return true;
}

// hard link named `name` among the link messages of a group's header; a missing name is an
// error only when `required`
bool find_link(H5File &f, const std::vector<Message> &msgs, const char *name, uint64_t &addr, bool required = true)
{
    bool links = false;
    for (const Message &m : msgs) {
        if (m.type == 0x11) return f.fail("symbol-table groups are not supported");
        if (m.type == 0x02) {   // link info: a defined fractal heap address means dense storage
            if (m.size < 2) return f.fail("link info truncated");
            const uint64_t hp = 2 + ((f.b[m.off + 1] & 1) ? 8 : 0);
            if (m.size < hp + 8) return f.fail("link info truncated");
            if (f.u(m.off + hp, 8) != kUndef) return f.fail("dense link storage is not supported");
        }
        if (m.type != 0x06) continue;
        links = true;
        uint64_t p = m.off;
        const uint64_t e = m.off + m.size;
        if (e - p < 2 || f.b[p] != 1) return f.fail("link message version is not 1");
        const unsigned lf = f.b[p + 1];
        p += 2;
        int ltype = 0;
        if (lf & 0x08) {
            if (p >= e) return f.fail("link message truncated");
            ltype = f.b[p++];
        }
        if (lf & 0x04) p += 8;   // creation order
        if (lf & 0x10) p += 1;   // name character set
        const int lb = 1 << (lf & 3);
        if (p > e || e - p < (uint64_t)lb) return f.fail("link message truncated");
        const uint64_t ln = f.u(p, lb);
        p += lb;
        if (ln > e - p) return f.fail("link name overruns its message");
        const bool match = ln == std::strlen(name) && std::memcmp(&f.b[p], name, ln) == 0;
        p += ln;
        if (!match) continue;
        if (ltype != 0) return f.fail(std::string("link '") + name + "' is not a hard link");
        if (e - p < 8) return f.fail("link message truncated");
        addr = f.u(p, 8);
        return true;
    }
    if (!f.err.empty() || !required) return false;
    return f.fail(std::string("no '") + name + "' in the group" + (links ? "" : " (no links)"));
}

// length of a rank-1 dataset of any type
bool dataset_length(H5File &f, uint64_t addr, uint64_t &n)
{
    std::vector<Message> msgs;
    if (!read_header(f, addr, msgs)) return false;
    for (const Message &m : msgs)
        if (m.type == 0x01) {
            const uint64_t o = m.off;
            if (m.size < 2 || f.b[o + 1] != 1) return f.fail("dimension is not rank 1");
            const uint64_t dpos = f.b[o] == 1 ? o + 8 : f.b[o] == 2 ? o + 4 : 0;
            if (!dpos || dpos + 8 > o + m.size) return f.fail("dimension dataspace unreadable");
            n = f.u(dpos, 8);
            return true;
        }
    return f.fail("dimension has no dataspace");
}

// a rank-1 float64 dataset: its length and (if out is set) its values
bool read_f64(H5File &f, uint64_t addr, const char *name, uint64_t &n, double *out, uint64_t cap)
{
    std::vector<Message> msgs;
    if (!read_header(f, addr, msgs)) return false;
    bool space = false, dtype = false, layout = false;
    uint64_t dim = 0, daddr = kUndef, dsize = 0;
    const std::string who = std::string("dataset '") + name + "': ";
    for (const Message &m : msgs) {
        const uint64_t o = m.off;
        if (m.type == 0x01) {   // dataspace
            if (m.size < 2) return f.fail(who + "dataspace truncated");
            const int ver = f.b[o], rank = f.b[o + 1];
            if (rank != 1) return f.fail(who + "rank is not 1");
            const uint64_t dpos = ver == 1 ? o + 8 : ver == 2 ? o + 4 : 0;
            if (!dpos) return f.fail(who + "dataspace version is not 1 or 2");
            if (dpos + 8 > o + m.size) return f.fail(who + "dataspace truncated");
            dim = f.u(dpos, 8);
            space = true;
        } else if (m.type == 0x03) {   // datatype
            if (m.size < 20) return f.fail(who + "datatype truncated");
            const unsigned cls = f.b[o] & 0x0f, bits0 = f.b[o + 1];
            const uint64_t size = f.u(o + 4, 4);
            if (cls != 1 || size != 8) return f.fail(who + "not a float64 dataset");
            if (bits0 & 0x41) return f.fail(who + "only little-endian floats are supported");
            // IEEE binary64: bit offset 0, precision 64, exponent at 52 (11 bits), mantissa at 0
            // (52 bits), bias 1023
            if (f.u(o + 8, 2) != 0 || f.u(o + 10, 2) != 64 || f.b[o + 12] != 52 || f.b[o + 13] != 11 ||
                f.b[o + 14] != 0 || f.b[o + 15] != 52 || f.u(o + 16, 4) != 1023)
                return f.fail(who + "not an IEEE binary64 layout");
            dtype = true;
        } else if (m.type == 0x08) {   // data layout
            if (m.size < 2) return f.fail(who + "layout truncated");
            const int ver = f.b[o], cls = f.b[o + 1];
            if (ver != 3 && ver != 4) return f.fail(who + "layout version is not 3 or 4");
            if (cls != 1) return f.fail(who + "only contiguous storage is supported");
            if (m.size < 18) return f.fail(who + "layout truncated");
            daddr = f.u(o + 2, 8);
            dsize = f.u(o + 10, 8);
            layout = true;
        } else if (m.type == 0x0b) {
            return f.fail(who + "filtered (compressed) datasets are not supported");
        }
    }
    if (!space || !dtype || !layout) return f.fail(who + "dataspace, datatype or layout missing");
    if (dim > (1ull << 40)) return f.fail(who + "length out of range");
    if (dsize < dim * 8) return f.fail(who + "storage smaller than the dataspace");
    n = dim;
    if (!out) return true;
    if (dim > cap) return f.fail(who + "more values than the output holds");
    if (dim == 0) return true;
    if (daddr == kUndef || daddr > ~0ull - f.base) return f.fail(who + "data never written");
    const uint64_t doff = daddr + f.base;
    if (!f.have(doff, dim * 8)) return f.fail(who + "data outside the file");
    for (uint64_t i = 0; i < dim; i++) {
        const uint64_t bits = f.u(doff + 8 * i, 8);
        std::memcpy(&out[i], &bits, 8);
    }
    return true;
}

bool read_nodes(H5File &f, const char *path, int &N, double *const out[5], int cap)
{
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return f.fail(std::string("cannot open ") + path);
    unsigned char chunk[1 << 16];
    size_t got;
    while ((got = std::fread(chunk, 1, sizeof(chunk), fp)) > 0) {
        f.b.insert(f.b.end(), chunk, chunk + got);
        if (f.b.size() > (1ull << 34)) {
            std::fclose(fp);
            return f.fail("file larger than 16 GiB");
        }
    }
    std::fclose(fp);
    static const unsigned char sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
    if (!f.have(0, 9) || std::memcmp(f.b.data(), sig, 8) != 0) return f.fail("not an HDF5 file");
    const int sv = f.b[8];
    uint64_t root = kUndef;
    if (sv == 0 || sv == 1) {
        if (!f.have(0, 96)) return f.fail("superblock truncated");
        if (f.b[13] != 8 || f.b[14] != 8) return f.fail("offsets or lengths are not 8 bytes");
        // version 1 inserts the indexed-storage K (2 bytes) and 2 reserved bytes before the
        // addresses: base address, free space, end of file, driver block, then the root entry
        const uint64_t addrs = sv == 0 ? 24 : 28;
        f.base = f.u(addrs, 8);
        const uint64_t entry = addrs + 32;   // root group symbol table entry
        root = f.u(entry + 8, 8);
    } else if (sv == 2 || sv == 3) {
        if (!f.have(0, 48)) return f.fail("superblock truncated");
        if (f.b[9] != 8 || f.b[10] != 8) return f.fail("offsets or lengths are not 8 bytes");
        f.base = f.u(12, 8);
        root = f.u(36, 8);
    } else {
        return f.fail("superblock version is not 0 to 3");
    }
    std::vector<Message> msgs;
    uint64_t group = 0, hid = 0, nhid = 0;
    if (!read_header(f, root, msgs) || !find_link(f, msgs, "nodeset", group)) return false;
    // the node count as the reference takes it: the length of the root dimension "hid"
    const bool have_hid = find_link(f, msgs, "hid", hid, false);
    if (!f.err.empty()) return false;
    if (have_hid && !dataset_length(f, hid, nhid)) return false;
    if (!read_header(f, group, msgs)) return false;
    static const char *names[5] = {"x", "y", "z", "lambda", "phi"};
    uint64_t addr[5], n0 = 0;
    for (int k = 0; k < 5; k++) {
        if (!find_link(f, msgs, names[k], addr[k])) return false;
        uint64_t n = 0;
        if (!read_f64(f, addr[k], names[k], n, nullptr, 0)) return false;
        if (k == 0) n0 = n;
        else if (n != n0) return f.fail("the five coordinate datasets differ in length");
    }
    if (n0 == 0 || n0 > 0x7fffffffull) return f.fail("node count out of range");
    if (have_hid && nhid != n0) return f.fail("dimension 'hid' differs from the coordinate length");
    N = (int)n0;
    if (!out[0]) return true;
    if (N > cap) return f.fail("more nodes than the output arrays hold");
    for (int k = 0; k < 5; k++) {
        uint64_t n = 0;
        if (out[k] && !read_f64(f, addr[k], names[k], n, out[k], (uint64_t)cap)) return false;
    }
    return true;
}

}  // namespace

extern "C" {

// serbf_read_nodeset: see serbf.h
int serbf_read_nodeset(const char *path, int *N, double *x, double *y, double *z, double *lambda, double *phi)
{
    if (!path || !N) {
        serbf::set_error("read_nodeset: null path or count");
        return SERBF_ERR_ARG;
    }
    const bool fill = x || y || z || lambda || phi;
    if (fill && !(x && y && z)) {
        serbf::set_error("read_nodeset: x, y and z are needed to read values");
        return SERBF_ERR_ARG;
    }
    H5File f;
    double *const out[5] = {x, y, z, lambda, phi};
    int n = 0;
    if (!read_nodes(f, path, n, out, fill ? *N : 0)) {
        serbf::set_error(std::string("read_nodeset: ") + path + ": " + f.err);
        return SERBF_ERR_ARG;
    }
    *N = n;
    return SERBF_OK;
}

}  // extern "C"
