/*
 * ncio.c -- netCDF classic files without the netCDF library.
 *
 * The reference writes its grid and state files with nc_create(...,
 * NC_CLOBBER), i.e. in the classic format (grid_generator/src/coordinator.c:354,
 * src/io/write_output.c:1428), and reads them with nc_open / nc_get_var_*
 * (src/io/set_grid_properties.c:32, src/io/set_initial_state.c:252).  This
 * file restates the published layout of that format (Unidata, "NetCDF
 * Classic Format Specification": CDF-1, and CDF-2 with 64-bit offsets) for
 * the subset those files use: fixed-size dimensions, one-dimensional NC_INT
 * and NC_DOUBLE variables, NC_CHAR attributes.  Everything on disk is
 * big-endian; names, attribute values and variable data are padded to four
 * bytes.
 *
 *   header   = 'C' 'D' 'F' version numrecs dim_list gatt_list var_list
 *   dim_list = ABSENT | 0x0A n [name length ...]
 *   gatt_list= ABSENT | 0x0C n [name type n values ...]
 *   var_list = ABSENT | 0x0B n [name ndims [dimid ...] vatt_list type vsize begin ...]
 *   ABSENT   = 0 0 (two 32-bit zeros); begin is 32-bit (CDF-1) or 64-bit (CDF-2)
 */
#define _FILE_OFFSET_BITS 64
#include <errno.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/game_files.h"

static char nc_err[512];

const char *game_nc_last_error(void) { return nc_err; }

#define NC_DIMENSION 0x0A
#define NC_VARIABLE 0x0B
#define NC_ATTRIBUTE 0x0C
#define NC_CHAR 2

static size_t pad4(size_t n) { return (n + 3) & ~(size_t)3; }
static int type_size(int t) { return t == GAME_NC_DOUBLE ? 8 : (t == GAME_NC_INT ? 4 : (t == NC_CHAR ? 1 : 0)); }

/* ------------------------------------------------------------------ */
/* writer                                                               */
/* ------------------------------------------------------------------ */
typedef struct { unsigned char *p; size_t n, cap; } buf_t;

static int put(buf_t *b, const void *src, size_t n)
{
    if (b->n + n > b->cap) {
        size_t c = b->cap ? 2*b->cap : 4096;
        while (c < b->n + n) c *= 2;
        unsigned char *q = realloc(b->p, c);
        if (!q) return 1;
        b->p = q;
        b->cap = c;
    }
    memcpy(b->p + b->n, src, n);
    b->n += n;
    return 0;
}
static int put32(buf_t *b, uint32_t v) { v = __builtin_bswap32(v); return put(b, &v, 4); }
static int put64(buf_t *b, uint64_t v) { v = __builtin_bswap64(v); return put(b, &v, 8); }
static int put_name(buf_t *b, const char *s)
{
    const size_t n = strlen(s);
    static const char zeros[4] = {0, 0, 0, 0};
    return put32(b, (uint32_t)n) || put(b, s, n) || put(b, zeros, pad4(n) - n);
}

/* the distinct dimensions of the variables, in order of first use */
static int dim_table(const game_nc_var_t *v, int n, const char **names, long long *lens, int *dimid)
{
    int nd = 0;
    for (int i = 0; i < n; ++i) {
        int k = 0;
        while (k < nd && strcmp(names[k], v[i].dim) != 0) ++k;
        if (k == nd) {
            names[nd] = v[i].dim;
            lens[nd] = v[i].len;
            ++nd;
        } else if (lens[k] != v[i].len) {
            snprintf(nc_err, sizeof nc_err, "dimension %s has lengths %lld and %lld", v[i].dim, lens[k], v[i].len);
            return -1;
        }
        dimid[i] = k;
    }
    return nd;
}

/* header of the file with the variables' data at `begin` (computed by the caller) */
static int header(buf_t *b, int version, const game_nc_var_t *v, int n, const char **dn, const long long *dl, int nd,
                  const int *dimid, const uint64_t *begin)
{
    const char magic[4] = {'C', 'D', 'F', (char)version};
    int rc = put(b, magic, 4) || put32(b, 0);
    rc = rc || put32(b, NC_DIMENSION) || put32(b, (uint32_t)nd);
    for (int k = 0; k < nd && !rc; ++k) rc = put_name(b, dn[k]) || put32(b, (uint32_t)dl[k]);
    rc = rc || put32(b, 0) || put32(b, 0);                     /* no global attributes */
    rc = rc || put32(b, NC_VARIABLE) || put32(b, (uint32_t)n);
    for (int i = 0; i < n && !rc; ++i) {
        rc = put_name(b, v[i].name) || put32(b, 1) || put32(b, (uint32_t)dimid[i]);
        if (v[i].units) {
            const size_t ul = strlen(v[i].units);
            static const char zeros[4] = {0, 0, 0, 0};
            rc = rc || put32(b, NC_ATTRIBUTE) || put32(b, 1) || put_name(b, "units") || put32(b, NC_CHAR)
                 || put32(b, (uint32_t)ul) || put(b, v[i].units, ul) || put(b, zeros, pad4(ul) - ul);
        } else rc = rc || put32(b, 0) || put32(b, 0);
        const uint64_t bytes = (uint64_t)v[i].len*(uint64_t)type_size(v[i].type);
        const uint64_t vs = pad4(bytes);
        rc = rc || put32(b, (uint32_t)v[i].type) || put32(b, vs >= 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)vs);
        rc = rc || (version == 1 ? put32(b, (uint32_t)begin[i]) : put64(b, begin[i]));
    }
    return rc;
}

int game_nc_write(const char *path, const game_nc_var_t *v, int n)
{
    nc_err[0] = 0;
    if (!path || !v || n < 1) { snprintf(nc_err, sizeof nc_err, "no variables"); return GAME_NC_EARG; }
    for (int i = 0; i < n; ++i)
        if (!v[i].name || !v[i].dim || v[i].len < 1 || !v[i].data || type_size(v[i].type) == 0
            || v[i].type == NC_CHAR) {
            snprintf(nc_err, sizeof nc_err, "variable %d: bad name, dimension, length, type or data", i);
            return GAME_NC_EARG;
        }
    const char **dn = malloc(sizeof(char *)*(size_t)n);
    long long *dl = malloc(sizeof(long long)*(size_t)n);
    int *dimid = malloc(sizeof(int)*(size_t)n);
    uint64_t *begin = calloc((size_t)n, sizeof(uint64_t));
    buf_t b = {0};
    int rc = GAME_NC_EIO;
    FILE *f = NULL;
    if (!dn || !dl || !dimid || !begin) goto out;
    const int nd = dim_table(v, n, dn, dl, dimid);
    if (nd < 0) { rc = GAME_NC_EARG; goto out; }
    /* the header's size does not depend on the offsets' values: lay it out
       once with zeros, then place the data after it */
    int version = 1;
    for (int pass = 0; pass < 2; ++pass) {
        b.n = 0;
        if (header(&b, version, v, n, dn, dl, nd, dimid, begin)) goto out;
        uint64_t off = b.n;
        for (int i = 0; i < n; ++i) {
            begin[i] = off;
            off += pad4((uint64_t)v[i].len*(uint64_t)type_size(v[i].type));
        }
        /* CDF-1 offsets are signed 32-bit: larger files take the 64-bit form */
        if (pass == 0 && begin[n - 1] > 0x7FFFFFFFull) { version = 2; continue; }
        b.n = 0;
        if (header(&b, version, v, n, dn, dl, nd, dimid, begin)) goto out;
        break;
    }
    f = fopen(path, "wb");
    if (!f) { snprintf(nc_err, sizeof nc_err, "%s: %s", path, strerror(errno)); goto out; }
    if (fwrite(b.p, 1, b.n, f) != b.n) goto werr;
    {
        enum { CHUNK = 1 << 16 };
        unsigned char *tmp = malloc(CHUNK*8);
        if (!tmp) goto werr;
        for (int i = 0; i < n; ++i) {
            const int ts = type_size(v[i].type);
            const unsigned char *src = v[i].data;
            for (long long k0 = 0; k0 < v[i].len; k0 += CHUNK) {
                const long long m = v[i].len - k0 < CHUNK ? v[i].len - k0 : CHUNK;
                for (long long k = 0; k < m; ++k) {
                    if (ts == 8) {
                        uint64_t x;
                        memcpy(&x, src + 8*(k0 + k), 8);
                        x = __builtin_bswap64(x);
                        memcpy(tmp + 8*k, &x, 8);
                    } else {
                        uint32_t x;
                        memcpy(&x, src + 4*(k0 + k), 4);
                        x = __builtin_bswap32(x);
                        memcpy(tmp + 4*k, &x, 4);
                    }
                }
                if (fwrite(tmp, (size_t)ts, (size_t)m, f) != (size_t)m) { free(tmp); goto werr; }
            }
            const size_t bytes = (size_t)v[i].len*(size_t)ts, padn = pad4(bytes) - bytes;
            static const unsigned char zeros[4] = {0, 0, 0, 0};
            if (padn && fwrite(zeros, 1, padn, f) != padn) { free(tmp); goto werr; }
        }
        free(tmp);
    }
    if (fclose(f) != 0) { f = NULL; goto werr; }
    f = NULL;
    rc = 0;
    goto out;
werr:
    snprintf(nc_err, sizeof nc_err, "%s: write failed", path);
out:
    if (f) fclose(f);
    free(b.p);
    free(dn); free(dl); free(dimid); free(begin);
    return rc;
}

/* ------------------------------------------------------------------ */
/* reader                                                               */
/* ------------------------------------------------------------------ */
typedef struct {
    char name[NC_NAME_MAX + 1];
    int type;
    long long len;           /* elements (the product of its dimensions) */
    uint64_t begin;
} nc_varinfo_t;

struct game_nc_file {
    FILE *f;
    int version;
    int nv;
    nc_varinfo_t *vars;
};

typedef struct { FILE *f; int bad; } rd_t;

static uint32_t get32(rd_t *r)
{
    uint32_t x = 0;
    if (fread(&x, 4, 1, r->f) != 1) r->bad = 1;
    return __builtin_bswap32(x);
}
static uint64_t get64(rd_t *r)
{
    uint64_t x = 0;
    if (fread(&x, 8, 1, r->f) != 1) r->bad = 1;
    return __builtin_bswap64(x);
}
static void get_name(rd_t *r, char *out)
{
    const uint32_t n = get32(r);
    if (r->bad || n > NC_NAME_MAX) { r->bad = 1; return; }
    char tmp[NC_NAME_MAX + 4];
    if (fread(tmp, 1, pad4(n), r->f) != pad4(n)) { r->bad = 1; return; }
    memcpy(out, tmp, n);
    out[n] = 0;
}
static void skip_atts(rd_t *r)
{
    const uint32_t tag = get32(r), n = get32(r);
    if (r->bad) return;
    if (tag == 0 && n == 0) return;
    if (tag != NC_ATTRIBUTE) { r->bad = 1; return; }
    for (uint32_t a = 0; a < n && !r->bad; ++a) {
        char name[NC_NAME_MAX + 1];
        get_name(r, name);
        const uint32_t t = get32(r), m = get32(r);
        const int ts = t == 1 || t == 2 ? 1 : (t == 3 ? 2 : (t == 4 || t == 5 ? 4 : (t == 6 ? 8 : 0)));
        if (r->bad || ts == 0) { r->bad = 1; return; }
        if (fseek(r->f, (long)pad4((size_t)m*(size_t)ts), SEEK_CUR) != 0) r->bad = 1;
    }
}

game_nc_file_t *game_nc_open(const char *path)
{
    nc_err[0] = 0;
    FILE *f = fopen(path, "rb");
    if (!f) { snprintf(nc_err, sizeof nc_err, "%s: %s", path, strerror(errno)); return NULL; }
    rd_t r = {f, 0};
    unsigned char magic[4];
    game_nc_file_t *h = NULL;
    long long *dlen = NULL;
    if (fread(magic, 1, 4, f) != 4 || magic[0] != 'C' || magic[1] != 'D' || magic[2] != 'F'
        || (magic[3] != 1 && magic[3] != 2)) {
        snprintf(nc_err, sizeof nc_err, "%s: not a netCDF classic (CDF-1/CDF-2) file", path);
        goto fail;
    }
    (void)get32(&r);                                    /* numrecs */
    {
        const uint32_t tag = get32(&r), nd = get32(&r);
        if (r.bad || (tag != NC_DIMENSION && !(tag == 0 && nd == 0))) goto bad;
        dlen = calloc(nd ? nd : 1, sizeof(long long));
        if (!dlen) goto bad;
        for (uint32_t k = 0; k < nd && !r.bad; ++k) {
            char name[NC_NAME_MAX + 1];
            get_name(&r, name);
            dlen[k] = get32(&r);
            if (dlen[k] == 0) {
                snprintf(nc_err, sizeof nc_err, "%s: record dimensions are not supported", path);
                goto fail;
            }
        }
        skip_atts(&r);                                  /* global attributes */
        const uint32_t vtag = get32(&r), nv = get32(&r);
        if (r.bad || (vtag != NC_VARIABLE && !(vtag == 0 && nv == 0))) goto bad;
        h = calloc(1, sizeof *h);
        if (!h) goto bad;
        h->vars = calloc(nv ? nv : 1, sizeof(nc_varinfo_t));
        if (!h->vars) goto bad;
        h->nv = (int)nv;
        h->version = magic[3];
        for (uint32_t i = 0; i < nv && !r.bad; ++i) {
            nc_varinfo_t *vi = &h->vars[i];
            get_name(&r, vi->name);
            const uint32_t ndims = get32(&r);
            long long len = 1;
            for (uint32_t k = 0; k < ndims && !r.bad; ++k) {
                const uint32_t id = get32(&r);
                if (id >= nd) { r.bad = 1; break; }
                len *= dlen[id];
            }
            skip_atts(&r);
            vi->type = (int)get32(&r);
            (void)get32(&r);                            /* vsize */
            vi->begin = h->version == 1 ? get32(&r) : get64(&r);
            vi->len = len;
        }
        if (r.bad) goto bad;
    }
    free(dlen);
    h->f = f;
    return h;
bad:
    snprintf(nc_err, sizeof nc_err, "%s: malformed netCDF header", path);
fail:
    if (h) { free(h->vars); free(h); }
    free(dlen);
    fclose(f);
    return NULL;
}

void game_nc_close(game_nc_file_t *h)
{
    if (!h) return;
    if (h->f) fclose(h->f);
    free(h->vars);
    free(h);
}

static const nc_varinfo_t *find(const game_nc_file_t *h, const char *name)
{
    for (int i = 0; i < h->nv; ++i) if (strcmp(h->vars[i].name, name) == 0) return &h->vars[i];
    return NULL;
}

int game_nc_inq(const game_nc_file_t *h, const char *name, int *type, long long *len)
{
    if (!h || !name) return GAME_NC_EARG;
    const nc_varinfo_t *vi = find(h, name);
    if (!vi) { snprintf(nc_err, sizeof nc_err, "no variable %s", name); return GAME_NC_ENOTVAR; }
    if (type) *type = vi->type;
    if (len) *len = vi->len;
    return 0;
}

int game_nc_get(game_nc_file_t *h, const char *name, int type, void *out, long long len)
{
    if (!h || !name || !out) return GAME_NC_EARG;
    const nc_varinfo_t *vi = find(h, name);
    if (!vi) { snprintf(nc_err, sizeof nc_err, "no variable %s", name); return GAME_NC_ENOTVAR; }
    if (vi->type != type || vi->len != len) {
        snprintf(nc_err, sizeof nc_err, "variable %s: type %d length %lld, asked for type %d length %lld", name,
                 vi->type, vi->len, type, len);
        return GAME_NC_EARG;
    }
    if (fseeko(h->f, (off_t)vi->begin, SEEK_SET) != 0) goto rerr;
    const int ts = type_size(type);
    if (fread(out, (size_t)ts, (size_t)len, h->f) != (size_t)len) goto rerr;
    unsigned char *p = out;
    if (ts == 8)
        for (long long k = 0; k < len; ++k) {
            uint64_t x;
            memcpy(&x, p + 8*k, 8);
            x = __builtin_bswap64(x);
            memcpy(p + 8*k, &x, 8);
        }
    else
        for (long long k = 0; k < len; ++k) {
            uint32_t x;
            memcpy(&x, p + 4*k, 4);
            x = __builtin_bswap32(x);
            memcpy(p + 4*k, &x, 4);
        }
    return 0;
rerr:
    snprintf(nc_err, sizeof nc_err, "variable %s: read failed", name);
    return GAME_NC_EIO;
}
