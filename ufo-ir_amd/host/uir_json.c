/*
 * uir_json.c -- ufo-core JSON graph runner for ufo-ir task chains.
 *
 * Reads the graph format of json-sample.json:1-56: "nodes" (plugin, name,
 * package, properties) and "edges" (from / to by node name).  The nodes must
 * form one chain: read -> one or more ufo-ir tasks -> write.  Task properties
 * are set through the task ABI by value type (numbers become float or uint by
 * the property's type, as GValue transforms do in ufo-core's loader);
 * object values with a "plugin" key become nested tasks (the method's
 * "projector", ASD-POCS's "df_minimizer").  Each frame is read from TIFF,
 * uploaded, passed through the chain with get_requisition / process, and
 * written back as a TIFF page.
 */
#define _GNU_SOURCE
#include <ctype.h>
#include <errno.h>
#include <glob.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>

#include "uir_internal.h"

/* ---- a small JSON tree ---- */

typedef enum { J_NULL, J_BOOL, J_NUM, J_STR, J_ARR, J_OBJ } jtype;

typedef struct jval {
    jtype t;
    double num;
    int b;
    char *str;
    struct jval **items; /* array items or object values */
    char **keys;         /* object keys */
    size_t n, cap;
} jval;

typedef struct {
    const char *s, *p;
    char err[160];
    int depth;
} jparser;

static void jfree(jval *v)
{
    if (!v)
        return;
    for (size_t i = 0; i < v->n; ++i) {
        jfree(v->items[i]);
        if (v->keys)
            free(v->keys[i]);
    }
    free(v->items);
    free(v->keys);
    free(v->str);
    free(v);
}

static void jerr(jparser *P, const char *what)
{
    if (P->err[0])
        return;
    int line = 1, col = 1;
    for (const char *q = P->s; q < P->p; ++q) {
        if (*q == '\n') {
            ++line;
            col = 1;
        } else {
            ++col;
        }
    }
    snprintf(P->err, sizeof P->err, "JSON %d:%d: %s", line, col, what);
}

static void ws(jparser *P)
{
    while (*P->p == ' ' || *P->p == '\t' || *P->p == '\n' || *P->p == '\r')
        ++P->p;
}

static jval *jnew(jtype t)
{
    jval *v = (jval *)calloc(1, sizeof *v);
    if (v)
        v->t = t;
    return v;
}

static int jpush(jval *v, char *key, jval *item)
{
    if (v->n == v->cap) {
        size_t c = v->cap ? 2 * v->cap : 8;
        jval **ni = (jval **)realloc(v->items, c * sizeof *ni);
        if (!ni)
            return -1;
        v->items = ni;
        if (v->t == J_OBJ) {
            char **nk = (char **)realloc(v->keys, c * sizeof *nk);
            if (!nk)
                return -1;
            v->keys = nk;
        }
        v->cap = c;
    }
    v->items[v->n] = item;
    if (v->t == J_OBJ)
        v->keys[v->n] = key;
    v->n++;
    return 0;
}

static void put_utf8(char **o, unsigned cp)
{
    char *q = *o;
    if (cp < 0x80) {
        *q++ = (char)cp;
    } else if (cp < 0x800) {
        *q++ = (char)(0xc0 | cp >> 6);
        *q++ = (char)(0x80 | (cp & 0x3f));
    } else {
        *q++ = (char)(0xe0 | cp >> 12);
        *q++ = (char)(0x80 | ((cp >> 6) & 0x3f));
        *q++ = (char)(0x80 | (cp & 0x3f));
    }
    *o = q;
}

static char *jstring(jparser *P)
{
    if (*P->p != '"') {
        jerr(P, "string expected");
        return NULL;
    }
    ++P->p;
    const char *start = P->p;
    size_t len = 0;
    for (const char *q = start; *q != '"'; ++q, ++len) {
        if (!*q) {
            P->p = q;
            jerr(P, "unterminated string");
            return NULL;
        }
        if (*q == '\\' && q[1])
            ++q;
    }
    char *out = (char *)malloc(len * 3 + 1), *o = out;
    if (!out) {
        jerr(P, "out of memory");
        return NULL;
    }
    while (*P->p != '"') {
        char c = *P->p++;
        if (c != '\\') {
            *o++ = c;
            continue;
        }
        c = *P->p++;
        switch (c) {
        case '"': *o++ = '"'; break;
        case '\\': *o++ = '\\'; break;
        case '/': *o++ = '/'; break;
        case 'b': *o++ = '\b'; break;
        case 'f': *o++ = '\f'; break;
        case 'n': *o++ = '\n'; break;
        case 'r': *o++ = '\r'; break;
        case 't': *o++ = '\t'; break;
        case 'u': {
            unsigned cp = 0;
            for (int k = 0; k < 4; ++k) {
                const char h = *P->p++;
                if (!isxdigit((unsigned char)h)) {
                    free(out);
                    jerr(P, "bad \\u escape");
                    return NULL;
                }
                cp = cp * 16 + (unsigned)(isdigit((unsigned char)h) ? h - '0' : (tolower((unsigned char)h) - 'a' + 10));
            }
            put_utf8(&o, cp);
            break;
        }
        default:
            free(out);
            jerr(P, "bad escape");
            return NULL;
        }
    }
    ++P->p;
    *o = 0;
    return out;
}

static jval *jparse(jparser *P);

static jval *jparse_obj(jparser *P)
{
    jval *v = jnew(J_OBJ);
    if (!v) {
        jerr(P, "out of memory");
        return NULL;
    }
    ++P->p;
    ws(P);
    if (*P->p == '}') {
        ++P->p;
        return v;
    }
    for (;;) {
        ws(P);
        char *k = jstring(P);
        if (!k)
            break;
        ws(P);
        if (*P->p != ':') {
            free(k);
            jerr(P, "':' expected");
            break;
        }
        ++P->p;
        jval *item = jparse(P);
        if (!item || jpush(v, k, item)) {
            free(k);
            jfree(item);
            jerr(P, "bad value");
            break;
        }
        ws(P);
        if (*P->p == ',') {
            ++P->p;
            continue;
        }
        if (*P->p == '}') {
            ++P->p;
            return v;
        }
        jerr(P, "',' or '}' expected");
        break;
    }
    jfree(v);
    return NULL;
}

static jval *jparse_arr(jparser *P)
{
    jval *v = jnew(J_ARR);
    if (!v) {
        jerr(P, "out of memory");
        return NULL;
    }
    ++P->p;
    ws(P);
    if (*P->p == ']') {
        ++P->p;
        return v;
    }
    for (;;) {
        jval *item = jparse(P);
        if (!item || jpush(v, NULL, item)) {
            jfree(item);
            jerr(P, "bad value");
            break;
        }
        ws(P);
        if (*P->p == ',') {
            ++P->p;
            continue;
        }
        if (*P->p == ']') {
            ++P->p;
            return v;
        }
        jerr(P, "',' or ']' expected");
        break;
    }
    jfree(v);
    return NULL;
}

static jval *jparse(jparser *P)
{
    ws(P);
    if (++P->depth > 64) {
        jerr(P, "nesting too deep");
        return NULL;
    }
    jval *v = NULL;
    const char c = *P->p;
    if (c == '{')
        v = jparse_obj(P);
    else if (c == '[')
        v = jparse_arr(P);
    else if (c == '"') {
        char *s = jstring(P);
        if (s && (v = jnew(J_STR)))
            v->str = s;
        else
            free(s);
    } else if (!strncmp(P->p, "true", 4) || !strncmp(P->p, "false", 5)) {
        if ((v = jnew(J_BOOL)))
            v->b = c == 't';
        P->p += c == 't' ? 4 : 5;
    } else if (!strncmp(P->p, "null", 4)) {
        v = jnew(J_NULL);
        P->p += 4;
    } else if (c == '-' || isdigit((unsigned char)c)) {
        char *end;
        errno = 0;
        const double d = strtod(P->p, &end);
        if (end == P->p || errno == ERANGE) {
            jerr(P, "bad number");
        } else if ((v = jnew(J_NUM))) {
            v->num = d;
            P->p = end;
        }
    } else {
        jerr(P, "value expected");
    }
    --P->depth;
    return v;
}

static const jval *jget(const jval *o, const char *key)
{
    if (!o || o->t != J_OBJ)
        return NULL;
    for (size_t i = 0; i < o->n; ++i)
        if (!strcmp(o->keys[i], key))
            return o->items[i];
    return NULL;
}

static const char *jstr(const jval *o, const char *key)
{
    const jval *v = jget(o, key);
    return v && v->t == J_STR ? v->str : NULL;
}

/* ---- task construction ---- */

static uir_status build_task(const jval *node, uir_task **out);

static uir_status set_props(uir_task *t, const jval *props)
{
    if (!props)
        return UIR_OK;
    if (props->t != J_OBJ)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "'properties' of '%s' is not an object", uir_task_plugin(t));
    for (size_t i = 0; i < props->n; ++i) {
        const char *k = props->keys[i];
        const jval *v = props->items[i];
        uir_prop_type pt;
        if (uir_task_property_type(t, k, &pt) != UIR_OK) {
            fprintf(stderr, "uir-json: warning: '%s' has no property '%s' (skipped)\n", uir_task_plugin(t), k);
            continue;
        }
        uir_status s = UIR_OK;
        switch (pt) {
        case UIR_PROP_FLOAT:
            if (v->t != J_NUM)
                return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "property '%s' wants a number", k);
            s = uir_task_set_float(t, k, (float)v->num);
            break;
        case UIR_PROP_UINT:
            if (v->t != J_NUM || v->num < 0 || v->num > 4294967295.0 || v->num != floor(v->num))
                return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "property '%s' wants an unsigned integer", k);
            s = uir_task_set_uint(t, k, (uint32_t)v->num);
            break;
        case UIR_PROP_BOOL:
            if (v->t != J_BOOL && v->t != J_NUM)
                return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "property '%s' wants a boolean", k);
            s = uir_task_set_bool(t, k, v->t == J_BOOL ? v->b : v->num != 0);
            break;
        case UIR_PROP_STRING:
            if (v->t != J_STR)
                return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "property '%s' wants a string", k);
            s = uir_task_set_string(t, k, v->str);
            break;
        case UIR_PROP_OBJECT: {
            if (v->t != J_OBJ || !jstr(v, "plugin"))
                return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "property '%s' wants a task object", k);
            uir_task *sub = NULL;
            if ((s = build_task(v, &sub)) != UIR_OK)
                return s;
            s = uir_task_set_object(t, k, sub);
            uir_task_unref(sub);
            break;
        }
        }
        if (s != UIR_OK)
            return s;
    }
    return UIR_OK;
}

static int is_ir_plugin(const char *plugin)
{
    static const char *ir[] = {"parallel-projector", "parallel_projector", "sirt", "sart", "asdpocs", "sbtv", NULL};
    for (int i = 0; ir[i]; ++i)
        if (!strcmp(plugin, ir[i]))
            return 1;
    return 0;
}

static uir_status build_task(const jval *node, uir_task **out)
{
    const char *plugin = jstr(node, "plugin");
    const char *package = jstr(node, "package");
    if (!plugin)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "node without 'plugin'");
    if (package && strcmp(package, "ir"))
        return uir_set_error(UIR_ERR_NOT_FOUND, "package '%s' is not available (only 'ir')", package);
    uir_task *t = uir_task_new(plugin);
    if (!t)
        return uir_set_error(UIR_ERR_NOT_FOUND, "no ir plugin '%s'", plugin);
    uir_status s = set_props(t, jget(node, "properties"));
    if (s != UIR_OK) {
        uir_task_unref(t);
        return s;
    }
    *out = t;
    return UIR_OK;
}

/* ---- read / write ---- */

static char *join(const char *base, const char *p)
{
    char *r = NULL;
    if (!base || !*base || p[0] == '/') {
        r = strdup(p);
    } else if (asprintf(&r, "%s/%s", base, p) < 0) {
        r = NULL;
    }
    return r;
}

typedef struct {
    char **files;
    size_t n_files;
    uint32_t *pages; /* per file */
    uint64_t total;
} inputs;

static void inputs_free(inputs *in)
{
    for (size_t i = 0; i < in->n_files; ++i)
        free(in->files[i]);
    free(in->files);
    free(in->pages);
    memset(in, 0, sizeof *in);
}

static int cmpstr(const void *a, const void *b) { return strcmp(*(char *const *)a, *(char *const *)b); }

static uir_status inputs_open(const char *path, inputs *in)
{
    memset(in, 0, sizeof *in);
    struct stat st;
    glob_t g;
    memset(&g, 0, sizeof g);
    int rc;
    if (stat(path, &st) == 0 && S_ISDIR(st.st_mode)) {
        char *pat = NULL;
        if (asprintf(&pat, "%s/*.tif*", path) < 0)
            return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "asprintf");
        rc = glob(pat, 0, NULL, &g);
        free(pat);
    } else {
        rc = glob(path, GLOB_NOCHECK, NULL, &g);
    }
    if (rc != 0 || g.gl_pathc == 0) {
        globfree(&g);
        return uir_set_error(UIR_ERR_NOT_FOUND, "read: nothing matches '%s'", path);
    }
    in->files = (char **)calloc(g.gl_pathc, sizeof(char *));
    in->pages = (uint32_t *)calloc(g.gl_pathc, sizeof(uint32_t));
    if (!in->files || !in->pages) {
        globfree(&g);
        inputs_free(in);
        return uir_set_error(UIR_ERR_OUT_OF_MEMORY, "calloc");
    }
    for (size_t i = 0; i < g.gl_pathc; ++i)
        in->files[in->n_files++] = strdup(g.gl_pathv[i]);
    globfree(&g);
    qsort(in->files, in->n_files, sizeof(char *), cmpstr);
    for (size_t i = 0; i < in->n_files; ++i) {
        uir_status s = uir_tiff_info(in->files[i], &in->pages[i], NULL, NULL);
        if (s != UIR_OK) {
            inputs_free(in);
            return s;
        }
        in->total += in->pages[i];
    }
    return UIR_OK;
}

/* Accepts one %[flags][width]{d,i,u,x,X} conversion (and %% escapes). */
static int frame_format(const char *f, int *has_conv)
{
    int convs = 0;
    for (const char *p = f; *p; ++p) {
        if (*p != '%')
            continue;
        if (p[1] == '%') {
            ++p;
            continue;
        }
        ++p;
        while (*p == '0' || *p == '-' || *p == '+' || *p == ' ')
            ++p;
        while (isdigit((unsigned char)*p))
            ++p;
        if (!(*p == 'd' || *p == 'i' || *p == 'u' || *p == 'x' || *p == 'X'))
            return -1;
        ++convs;
    }
    if (convs > 1)
        return -1;
    *has_conv = convs;
    return 0;
}

static void mkdirs_for(const char *file)
{
    char *d = strdup(file);
    if (!d)
        return;
    for (char *p = d + 1; *p; ++p)
        if (*p == '/') {
            *p = 0;
            mkdir(d, 0775);
            *p = '/';
        }
    free(d);
}

/* ---- runner ---- */

typedef struct {
    const jval *node;
    const char *name;
    int next, has_in;
} gnode;

uir_status uir_json_run(const char *json_text, const char *base_dir, int device, uir_json_report *report)
{
    if (!json_text)
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "NULL JSON");
    jparser P = {json_text, json_text, {0}, 0};
    jval *root = jparse(&P);
    if (root) {
        ws(&P);
        if (*P.p)
            jerr(&P, "trailing characters");
    }
    if (P.err[0]) {
        jfree(root);
        return uir_set_error(UIR_ERR_INVALID_ARGUMENT, "%s", P.err);
    }
    uir_status s = UIR_OK;
    gnode *g = NULL;
    uir_task **chain = NULL;
    size_t n_chain = 0;
    uir_resources *res = NULL;
    inputs in = {0};
    float *h_in = NULL, *h_out = NULL;
    void **dbuf = NULL;
    size_t *dcap = NULL;
    char *rpath = NULL, *wpath = NULL, *fname = NULL;
    size_t cap_in = 0, cap_out = 0;
    const jval *nodes = jget(root, "nodes"), *edges = jget(root, "edges");
    if (!nodes || nodes->t != J_ARR || !edges || edges->t != J_ARR) {
        s = uir_set_error(UIR_ERR_INVALID_ARGUMENT, "graph needs 'nodes' and 'edges' arrays");
        goto out;
    }
    g = (gnode *)calloc(nodes->n ? nodes->n : 1, sizeof *g);
    if (!g) {
        s = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "calloc");
        goto out;
    }
    for (size_t i = 0; i < nodes->n; ++i) {
        g[i].node = nodes->items[i];
        g[i].name = jstr(g[i].node, "name");
        g[i].next = -1;
        if (!g[i].name || !jstr(g[i].node, "plugin")) {
            s = uir_set_error(UIR_ERR_INVALID_ARGUMENT, "node %zu needs 'name' and 'plugin'", i);
            goto out;
        }
        for (size_t j = 0; j < i; ++j)
            if (!strcmp(g[j].name, g[i].name)) {
                s = uir_set_error(UIR_ERR_INVALID_ARGUMENT, "two nodes named '%s'", g[i].name);
                goto out;
            }
    }
    for (size_t e = 0; e < edges->n; ++e) {
        const char *from = jstr(jget(edges->items[e], "from"), "name");
        const char *to = jstr(jget(edges->items[e], "to"), "name");
        int a = -1, b = -1;
        for (size_t i = 0; i < nodes->n; ++i) {
            if (from && !strcmp(g[i].name, from)) a = (int)i;
            if (to && !strcmp(g[i].name, to)) b = (int)i;
        }
        if (a < 0 || b < 0) {
            s = uir_set_error(UIR_ERR_INVALID_ARGUMENT, "edge %zu names an unknown node", e);
            goto out;
        }
        if (g[a].next >= 0 || g[b].has_in) {
            s = uir_set_error(UIR_ERR_UNSUPPORTED, "only a single chain of nodes is supported (node '%s' or '%s')",
                              g[a].name, g[b].name);
            goto out;
        }
        g[a].next = b;
        g[b].has_in = 1;
    }
    int head = -1;
    for (size_t i = 0; i < nodes->n; ++i)
        if (!g[i].has_in) {
            if (head >= 0) {
                s = uir_set_error(UIR_ERR_UNSUPPORTED, "the graph has more than one source node");
                goto out;
            }
            head = (int)i;
        }
    if (head < 0 || strcmp(jstr(g[head].node, "plugin"), "read")) {
        s = uir_set_error(UIR_ERR_UNSUPPORTED, "the chain must start with a 'read' node");
        goto out;
    }
    /* walk: read -> ir tasks -> write */
    chain = (uir_task **)calloc(nodes->n, sizeof *chain);
    if (!chain) {
        s = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "calloc");
        goto out;
    }
    int cur = g[head].next, tail = -1;
    size_t steps = 0;
    while (cur >= 0 && steps++ <= nodes->n) {
        const char *plugin = jstr(g[cur].node, "plugin");
        if (!strcmp(plugin, "write")) {
            tail = cur;
            break;
        }
        if (!is_ir_plugin(plugin)) {
            s = uir_set_error(UIR_ERR_NOT_FOUND, "plugin '%s' is not available (read, write and ir tasks are)", plugin);
            goto out;
        }
        if ((s = build_task(g[cur].node, &chain[n_chain])) != UIR_OK)
            goto out;
        ++n_chain;
        cur = g[cur].next;
    }
    if (tail < 0 || g[tail].next >= 0 || n_chain == 0) {
        s = uir_set_error(UIR_ERR_UNSUPPORTED, "the chain must be read -> ir task(s) -> write");
        goto out;
    }
    for (size_t i = 0; i < nodes->n; ++i) {
        int reached = (int)i == head;
        for (int c = g[head].next; c >= 0 && !reached; c = g[c].next)
            reached = c == (int)i;
        if (!reached) {
            s = uir_set_error(UIR_ERR_UNSUPPORTED, "node '%s' is not on the chain", g[i].name);
            goto out;
        }
    }
    /* read / write properties */
    const jval *rp = jget(g[head].node, "properties"), *wp = jget(g[tail].node, "properties");
    const char *rpath_rel = jstr(rp, "path");
    const char *wfmt = jstr(wp, "filename");
    if (!rpath_rel) {
        s = uir_set_error(UIR_ERR_INVALID_ARGUMENT, "read needs a 'path'");
        goto out;
    }
    if (!wfmt)
        wfmt = "./output-%05i.tif";
    int has_conv = 0;
    if (frame_format(wfmt, &has_conv)) {
        s = uir_set_error(UIR_ERR_INVALID_ARGUMENT, "write: 'filename' may hold one integer conversion only");
        goto out;
    }
    uint64_t start = 0, number = UINT64_MAX;
    for (size_t i = 0; rp && rp->t == J_OBJ && i < rp->n; ++i) {
        const char *k = rp->keys[i];
        const jval *v = rp->items[i];
        if (!strcmp(k, "path"))
            continue;
        if ((!strcmp(k, "start") || !strcmp(k, "number")) && v->t == J_NUM && v->num >= 0) {
            *(!strcmp(k, "start") ? &start : &number) = (uint64_t)v->num;
            continue;
        }
        fprintf(stderr, "uir-json: warning: read: property '%s' is not supported (skipped)\n", k);
    }
    for (size_t i = 0; wp && wp->t == J_OBJ && i < wp->n; ++i)
        if (strcmp(wp->keys[i], "filename"))
            fprintf(stderr, "uir-json: warning: write: property '%s' is not supported (skipped)\n", wp->keys[i]);
    rpath = join(base_dir, rpath_rel);
    wpath = strdup(wfmt); /* formatted per frame, then joined with base_dir */
    if (!rpath || !wpath) {
        s = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "path");
        goto out;
    }
    if ((s = inputs_open(rpath, &in)) != UIR_OK)
        goto out;
    /* device side */
    if ((s = uir_resources_new(device, &res)) != UIR_OK)
        goto out;
    for (size_t i = 0; i < n_chain; ++i)
        if ((s = uir_task_setup(chain[i], res)) != UIR_OK)
            goto out;
    dbuf = (void **)calloc(n_chain + 1, sizeof(void *));
    dcap = (size_t *)calloc(n_chain + 1, sizeof(size_t));
    if (!dbuf || !dcap) {
        s = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "calloc");
        goto out;
    }
    const double t0 = uir_clock();
    uint64_t frame = 0, written = 0;
    for (size_t fi = 0; fi < in.n_files && s == UIR_OK; ++fi)
        for (uint32_t pg = 0; pg < in.pages[fi] && s == UIR_OK; ++pg, ++frame) {
            if (frame < start)
                continue;
            if (written >= number)
                break;
            uint32_t w = 0, h = 0;
            if ((s = uir_tiff_read(in.files[fi], pg, NULL, 0, &w, &h)) != UIR_OK)
                break;
            const size_t nin = (size_t)w * h;
            if (nin > cap_in) {
                free(h_in);
                h_in = (float *)malloc(nin * 4);
                cap_in = h_in ? nin : 0;
                if (!h_in) {
                    s = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "malloc");
                    break;
                }
            }
            if ((s = uir_tiff_read(in.files[fi], pg, h_in, cap_in, &w, &h)) != UIR_OK)
                break;
            if (nin * 4 > dcap[0]) {
                uir_device_free(dbuf[0]);
                dbuf[0] = NULL;
                dcap[0] = 0;
                if ((s = uir_device_malloc(nin * 4, &dbuf[0])) != UIR_OK)
                    break;
                dcap[0] = nin * 4;
            }
            if ((s = uir_memcpy_htod(dbuf[0], h_in, nin * 4, res->stream)) != UIR_OK)
                break;
            uir_buffer cur_buf = {dbuf[0], {w, h}, UIR_DTYPE_F32};
            for (size_t k = 0; k < n_chain && s == UIR_OK; ++k) {
                uir_requisition rq;
                if ((s = uir_task_get_requisition(chain[k], &cur_buf, &rq)) != UIR_OK)
                    break;
                const size_t nout = (size_t)rq.dims[0] * rq.dims[1];
                if (nout * 4 > dcap[k + 1]) {
                    uir_device_free(dbuf[k + 1]);
                    dbuf[k + 1] = NULL;
                    dcap[k + 1] = 0;
                    if ((s = uir_device_malloc(nout * 4, &dbuf[k + 1])) != UIR_OK)
                        break;
                    dcap[k + 1] = nout * 4;
                }
                uir_buffer ob = {dbuf[k + 1], {(uint32_t)rq.dims[0], (uint32_t)rq.dims[1]}, UIR_DTYPE_F32};
                if ((s = uir_task_process(chain[k], &cur_buf, &ob)) != UIR_OK)
                    break;
                cur_buf = ob;
            }
            if (s != UIR_OK)
                break;
            const size_t nout = (size_t)cur_buf.dims[0] * cur_buf.dims[1];
            if (nout > cap_out) {
                free(h_out);
                h_out = (float *)malloc(nout * 4);
                cap_out = h_out ? nout : 0;
                if (!h_out) {
                    s = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "malloc");
                    break;
                }
            }
            if ((s = uir_memcpy_dtoh(h_out, cur_buf.data, nout * 4, res->stream)) != UIR_OK ||
                (s = uir_stream_synchronize(res->stream)) != UIR_OK)
                break;
            free(fname);
            fname = NULL;
            {
                char *rel = NULL;
                if (has_conv) {
                    if (asprintf(&rel, wpath, (int)written) < 0) /* wpath holds one checked conversion */
                        rel = NULL;
                } else {
                    rel = strdup(wpath);
                }
                if (rel)
                    fname = join(base_dir, rel);
                free(rel);
            }
            if (!fname) {
                s = uir_set_error(UIR_ERR_OUT_OF_MEMORY, "file name");
                break;
            }
            mkdirs_for(fname);
            s = uir_tiff_write(fname, h_out, cur_buf.dims[0], cur_buf.dims[1], !has_conv && written > 0);
            ++written;
        }
    if (report && s == UIR_OK) {
        report->frames = (uint32_t)written;
        report->tasks = (uint32_t)n_chain;
        report->seconds = uir_clock() - t0;
    }
out:
    if (dbuf)
        for (size_t k = 0; k <= n_chain; ++k)
            uir_device_free(dbuf[k]);
    free(dbuf);
    free(dcap);
    free(h_in);
    free(h_out);
    free(fname);
    free(rpath);
    free(wpath);
    inputs_free(&in);
    for (size_t i = 0; i < n_chain; ++i)
        uir_task_unref(chain[i]);
    free(chain);
    uir_resources_free(res);
    free(g);
    jfree(root);
    return s;
}
