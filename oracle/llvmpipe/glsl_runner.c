/*
 * glsl_runner — runs the reference's own GLSL compute shaders on Mesa llvmpipe
 * (host CPU cores) and dumps every buffer the Barnes-Hut step produces.
 *
 * This is test infrastructure: it produces golden vectors for the CPU oracle
 * and for the HIP path. It is built and run only in the development container
 * (it reads the shader sources from the reference checkout at run time); its
 * outputs are packed into small fixtures under tests/golden/.
 *
 * What it replicates from the Java host (reference files cited):
 *  - shader assembly: recursive `#include "x"` expansion
 *    (gpu/GLSLShader.java:102-121), `#define COMPUTE_*` inserted after the
 *    `#version`/`#extension` preamble (gpu/ComputeShader.java:53-70), removal of
 *    every `//For render shaders: ... //End for render shaders` block
 *    (gpu/ComputeShader.java:27). One addition: `#extension
 *    GL_ARB_gpu_shader_int64 : enable` is placed in the preamble, because Mesa
 *    does not expose GL_NV_gpu_shader5 (which the shaders request for uint64_t).
 *  - buffer sizes per binding (gpu/GPU.java:199-341), `sim` packing
 *    (gpu/GPU.java:1096-1143), upload of bodies into BodiesIn (gpu/GPU.java:1166-1196)
 *  - per-program uniform lists and work-group counts (gpu/GPU.java:402-832,1149)
 *  - the step's dispatch order and buffer-name swaps
 *    (simulation/BarnesHut.java:99-164,358-448,474-535; gpu/GPU.java:1060-1086)
 *  - glMemoryBarrier(GL_SHADER_STORAGE_BARRIER_BIT) after every dispatch
 *    (gpu/ComputeProgram.java:51-55)
 *
 * The GL context is created surfaceless through the swrast DRI driver's public
 * interface (GL/internal/dri_interface.h); GL entry points come from
 * libglapi's _glapi_get_proc_address.
 *
 * Input file (little-endian):  struct run_header, then N*8 float32 bodies
 * (posMass, velDensity). Output: raw buffer dumps "<out>/s<step>_<stage>_<buf>.bin"
 * and a JSON timing summary on stdout.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include <time.h>
#include <GL/gl.h>
#include <GL/glext.h>
#include <GL/internal/dri_interface.h>

struct run_header {
    uint32_t magic;          /* 'BHIN' = 0x4E494842 */
    uint32_t n;
    uint32_t steps;
    uint32_t merging;        /* mergingCollisionOrNeither: 0 none, 1 merge, 2 collision */
    uint32_t dynamic;        /* staticOrDynamic: 0 static, 1 dynamic */
    uint32_t wrap;           /* wrapAround */
    uint32_t flags;          /* bit0 dump stages, bit1 acceleration extraction, bit2 timing */
    uint32_t dump_from;      /* first step whose stages are dumped */
    float theta, dt, softening, elasticity;
    float mass, density, len, time, camera_scale;
    float bounds[6];
};

/* ------------------------------------------------------------------ GL */
typedef void *(*gpa_t)(const char *);
static gpa_t gpa;
#define GLF(ret, name, args) static ret (*p##name) args;
GLF(GLuint, glCreateShader, (GLenum))
GLF(void, glShaderSource, (GLuint, GLsizei, const GLchar *const *, const GLint *))
GLF(void, glCompileShader, (GLuint))
GLF(void, glGetShaderiv, (GLuint, GLenum, GLint *))
GLF(void, glGetShaderInfoLog, (GLuint, GLsizei, GLsizei *, GLchar *))
GLF(GLuint, glCreateProgram, (void))
GLF(void, glAttachShader, (GLuint, GLuint))
GLF(void, glLinkProgram, (GLuint))
GLF(void, glGetProgramiv, (GLuint, GLenum, GLint *))
GLF(void, glGetProgramInfoLog, (GLuint, GLsizei, GLsizei *, GLchar *))
GLF(void, glUseProgram, (GLuint))
GLF(GLint, glGetUniformLocation, (GLuint, const GLchar *))
GLF(void, glUniform1ui, (GLint, GLuint))
GLF(void, glUniform1f, (GLint, GLfloat))
GLF(void, glGenBuffers, (GLsizei, GLuint *))
GLF(void, glBindBuffer, (GLenum, GLuint))
GLF(void, glBufferData, (GLenum, GLsizeiptr, const void *, GLenum))
GLF(void, glBufferSubData, (GLenum, GLintptr, GLsizeiptr, const void *))
GLF(void, glGetBufferSubData, (GLenum, GLintptr, GLsizeiptr, void *))
GLF(void, glBindBufferBase, (GLenum, GLuint, GLuint))
GLF(void, glDispatchCompute, (GLuint, GLuint, GLuint))
GLF(void, glMemoryBarrier, (GLbitfield))
GLF(void, glFinish, (void))
GLF(GLenum, glGetError, (void))
GLF(void, glGetIntegerv, (GLenum, GLint *))
GLF(const GLubyte *, glGetString, (GLenum))
GLF(GLuint, glGetProgramResourceIndex, (GLuint, GLenum, const GLchar *))
GLF(void, glGetProgramResourceiv, (GLuint, GLenum, GLuint, GLsizei, const GLenum *, GLsizei, GLsizei *, GLint *))

static void load_gl(void) {
#define L(name) do { p##name = gpa(#name); if (!p##name) { fprintf(stderr, "missing %s\n", #name); exit(2); } } while (0)
    L(glCreateShader); L(glShaderSource); L(glCompileShader); L(glGetShaderiv);
    L(glGetShaderInfoLog); L(glCreateProgram); L(glAttachShader); L(glLinkProgram);
    L(glGetProgramiv); L(glGetProgramInfoLog); L(glUseProgram); L(glGetUniformLocation);
    L(glUniform1ui); L(glUniform1f); L(glGenBuffers); L(glBindBuffer); L(glBufferData);
    L(glBufferSubData); L(glGetBufferSubData); L(glBindBufferBase); L(glDispatchCompute);
    L(glMemoryBarrier); L(glFinish); L(glGetError); L(glGetIntegerv); L(glGetString);
    L(glGetProgramResourceIndex); L(glGetProgramResourceiv);
#undef L
}

static void dgi(__DRIdrawable *d, int *x, int *y, int *w, int *h, void *p) {
    (void)d; (void)p; *x = *y = 0; *w = *h = 1;
}
static void dpi(__DRIdrawable *d, int op, int x, int y, int w, int h, char *data, void *p) {
    (void)d; (void)op; (void)x; (void)y; (void)w; (void)h; (void)data; (void)p;
}
static void dgim(__DRIdrawable *d, int x, int y, int w, int h, char *data, void *p) {
    (void)d; (void)x; (void)y; (void)w; (void)h; (void)data; (void)p;
}
static const __DRIswrastLoaderExtension swrast_loader = {{__DRI_SWRAST_LOADER, 1}, dgi, dpi, dgim};
static const __DRIextension *loader_exts[] = {&swrast_loader.base, NULL};

static void create_context(void) {
    const char *path = getenv("BH_SWRAST_DRI");
    if (!path) path = "/usr/lib/x86_64-linux-gnu/dri/swrast_dri.so";
    void *drv = dlopen(path, RTLD_NOW | RTLD_GLOBAL);
    if (!drv) { fprintf(stderr, "dlopen %s: %s\n", path, dlerror()); exit(2); }
    const __DRIextension **(*get)(void) = (const __DRIextension **(*)(void))dlsym(drv, "__driDriverGetExtensions_swrast");
    if (!get) { fprintf(stderr, "no __driDriverGetExtensions_swrast\n"); exit(2); }
    const __DRIextension **ext = get();
    const __DRIcoreExtension *core = NULL;
    const __DRIswrastExtension *sw = NULL;
    for (int i = 0; ext[i]; i++) {
        if (!strcmp(ext[i]->name, __DRI_CORE)) core = (const __DRIcoreExtension *)ext[i];
        if (!strcmp(ext[i]->name, __DRI_SWRAST)) sw = (const __DRIswrastExtension *)ext[i];
    }
    if (!core || !sw || sw->base.version < 4) { fprintf(stderr, "swrast extensions missing\n"); exit(2); }
    const __DRIconfig **configs = NULL;
    __DRIscreen *scr = sw->createNewScreen2(0, loader_exts, ext, &configs, NULL);
    if (!scr) { fprintf(stderr, "createNewScreen2 failed\n"); exit(2); }
    uint32_t attribs[] = {__DRI_CTX_ATTRIB_MAJOR_VERSION, 4, __DRI_CTX_ATTRIB_MINOR_VERSION, 3};
    unsigned err = 0;
    __DRIcontext *ctx = sw->createContextAttribs(scr, __DRI_API_OPENGL_CORE, NULL, NULL, 2, attribs, &err, NULL);
    if (!ctx) { fprintf(stderr, "createContextAttribs failed (%u)\n", err); exit(2); }
    if (!core->bindContext(ctx, NULL, NULL)) { fprintf(stderr, "bindContext failed\n"); exit(2); }
    gpa = (gpa_t)dlsym(RTLD_DEFAULT, "_glapi_get_proc_address");
    if (!gpa) { fprintf(stderr, "no _glapi_get_proc_address\n"); exit(2); }
    load_gl();
}

static void check_gl(const char *what) {
    GLenum e = pglGetError();
    if (e != GL_NO_ERROR) { fprintf(stderr, "GL error 0x%x after %s\n", e, what); exit(3); }
}

/* ------------------------------------------------------------ sources */
static char *shader_root;

static char *read_file(const char *path) {
    FILE *f = fopen(path, "rb");
    if (!f) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
    fseek(f, 0, SEEK_END); long n = ftell(f); fseek(f, 0, SEEK_SET);
    char *b = malloc(n + 1);
    if (fread(b, 1, n, f) != (size_t)n) { fprintf(stderr, "short read %s\n", path); exit(2); }
    b[n] = 0; fclose(f);
    return b;
}

typedef struct { char *p; size_t n, cap; } sbuf;
static void sb_add(sbuf *s, const char *t, size_t n) {
    if (s->n + n + 1 > s->cap) { s->cap = (s->n + n + 1) * 2; s->p = realloc(s->p, s->cap); }
    memcpy(s->p + s->n, t, n); s->n += n; s->p[s->n] = 0;
}

/* recursive #include "x" expansion, one output line per input line */
static void expand(sbuf *out, const char *rel) {
    char path[4096];
    snprintf(path, sizeof path, "%s/%s", shader_root, rel);
    char *src = read_file(path);
    char *line = src;
    for (;;) {
        char *nl = strchr(line, '\n');
        size_t len = nl ? (size_t)(nl - line) : strlen(line);
        const char *t = line;
        while (t < line + len && (*t == ' ' || *t == '\t' || *t == '\r')) t++;
        int handled = 0;
        if (!strncmp(t, "#include\"", 9) || !strncmp(t, "#include \"", 10)) {
            const char *a = memchr(t, '"', line + len - t);
            const char *b = NULL;
            for (const char *q = line + len - 1; q > a; q--) if (*q == '"') { b = q; break; }
            if (a && b && b > a + 1) {
                char inc[1024];
                size_t m = (size_t)(b - a - 1);
                memcpy(inc, a + 1, m); inc[m] = 0;
                expand(out, inc);
                sb_add(out, "\n", 1);
                handled = 1;
            }
        }
        if (!handled) { sb_add(out, line, len); sb_add(out, "\n", 1); }
        if (!nl) break;
        line = nl + 1;
    }
    free(src);
}

static char *strip_render_blocks(const char *s) {
    const char *open = "//For render shaders:", *close = "//End for render shaders";
    sbuf o = {0};
    for (;;) {
        const char *a = strstr(s, open);
        if (!a) { sb_add(&o, s, strlen(s)); break; }
        const char *b = strstr(a, close);
        if (!b) { sb_add(&o, s, strlen(s)); break; }
        sb_add(&o, s, (size_t)(a - s));
        sb_add(&o, "//Removed render shader code here", 33);
        s = b + strlen(close);
    }
    return o.p;
}

static char *program_source(const char *define) {
    sbuf all = {0};
    expand(&all, "compute/bh_main.comp");
    /* insert after the #version/#extension preamble */
    size_t pos = 0, ins = 0;
    while (pos < all.n) {
        char *nl = strchr(all.p + pos, '\n');
        size_t end = nl ? (size_t)(nl - all.p) : all.n;
        const char *t = all.p + pos;
        while (*t == ' ' || *t == '\t') t++;
        if (!strncmp(t, "#version", 8) || !strncmp(t, "#extension", 10)) { ins = end + 1; pos = end + 1; continue; }
        break;
    }
    sbuf s = {0};
    sb_add(&s, all.p, ins);
    const char *int64 = "#extension GL_ARB_gpu_shader_int64 : enable\n";
    sb_add(&s, int64, strlen(int64));
    char def[256];
    snprintf(def, sizeof def, "#define %s\n", define);
    sb_add(&s, def, strlen(def));
    sb_add(&s, all.p + ins, all.n - ins);
    free(all.p);
    char *r = strip_render_blocks(s.p);
    free(s.p);
    return r;
}

static GLuint build_program(const char *define) {
    char *src = program_source(define);
    GLuint sh = pglCreateShader(GL_COMPUTE_SHADER);
    const char *srcs[1] = {src};
    pglShaderSource(sh, 1, srcs, NULL);
    pglCompileShader(sh);
    GLint ok = 0;
    pglGetShaderiv(sh, GL_COMPILE_STATUS, &ok);
    if (!ok) {
        char log[16384]; GLsizei n = 0;
        pglGetShaderInfoLog(sh, sizeof log, &n, log);
        fprintf(stderr, "compile %s failed:\n%s\n", define, log);
        exit(4);
    }
    GLuint prog = pglCreateProgram();
    pglAttachShader(prog, sh);
    pglLinkProgram(prog);
    pglGetProgramiv(prog, GL_LINK_STATUS, &ok);
    if (!ok) {
        char log[16384]; GLsizei n = 0;
        pglGetProgramInfoLog(prog, sizeof log, &n, log);
        fprintf(stderr, "link %s failed:\n%s\n", define, log);
        exit(4);
    }
    free(src);
    return prog;
}

/* ------------------------------------------------------------ state */
enum { B_LEAF = 0, B_INTERNAL, B_SIM, B_BODIES_IN, B_BODIES_OUT, B_MORTON_IN, B_MORTON_OUT,
       B_INDEX_IN, B_INDEX_OUT, B_WQ_IN, B_WQ_OUT, B_HIST, B_SCANNED, B_TOTALS, B_MERGE, B_LOCKS, B_COUNT };
static const char *bname[B_COUNT] = {"leaf", "internal", "sim", "bodies_in", "bodies_out", "morton_in",
    "morton_out", "index_in", "index_out", "wq_in", "wq_out", "hist", "scanned", "totals", "merge", "locks"};

static GLuint buf[B_COUNT];       /* current GL name bound at each binding */
static size_t bsize[B_COUNT];
static struct run_header H;
static uint32_t N, G;
static const char *outdir;

enum { P_INIT, P_UPDATE, P_DEAD_COUNT, P_DEAD_SCAN, P_DEAD_SCATTER, P_AABB_REPOP, P_AABB_COLLAPSE,
       P_ENCODE, P_HIST, P_BUCKET_SCAN, P_GLOBAL_SCAN, P_SCATTER, P_TREE_BUILD, P_INIT_LEAVES,
       P_PROPAGATE, P_FORCE, P_MERGE, P_COUNT };
static const char *pdef[P_COUNT] = {"COMPUTE_INIT", "COMPUTE_UPDATE", "COMPUTE_DEAD_COUNT",
    "COMPUTE_DEAD_EXCLUSIVE_SCAN", "COMPUTE_DEAD_SCATTER", "COMPUTE_MORTON_AABB_REPOPULATE",
    "COMPUTE_MORTON_AABB_COLLAPSE", "COMPUTE_MORTON_ENCODE", "COMPUTE_RADIX_HIST",
    "COMPUTE_RADIX_BUCKET_SCAN", "COMPUTE_RADIX_GLOBAL_SCAN", "COMPUTE_RADIX_SCATTER",
    "COMPUTE_TREE_BUILD", "COMPUTE_TREE_INIT_LEAVES", "COMPUTE_TREE_PROPAGATE_NODES",
    "COMPUTE_FORCE_COMPUTE", "COMPUTE_MERGE_BODIES"};
static GLuint prog[P_COUNT];

/* host-side uniform values (gpu/GPU.java:402-498) */
static uint32_t u_pass_shift, u_reset_or_decrement;

static void swap_names(int a, int b) { GLuint t = buf[a]; buf[a] = buf[b]; buf[b] = t; }

static void set_f(GLuint p, const char *n, float v) { GLint l = pglGetUniformLocation(p, n); if (l >= 0) pglUniform1f(l, v); }
static void set_u(GLuint p, const char *n, uint32_t v) { GLint l = pglGetUniformLocation(p, n); if (l >= 0) pglUniform1ui(l, v); }

static double now_ms(void) { struct timespec t; clock_gettime(CLOCK_MONOTONIC, &t); return t.tv_sec * 1e3 + t.tv_nsec / 1e6; }

static void dispatch(int p, uint32_t groups) {
    pglUseProgram(prog[p]);
    /* uniforms: per-program lists of gpu/GPU.java:504-832 (unused ones are optimised away) */
    switch (p) {
    case P_UPDATE:
        set_u(prog[p], "resetValuesOrDecrementDeadBodies", u_reset_or_decrement);
        set_f(prog[p], "mass", H.mass); set_f(prog[p], "density", H.density);
        set_f(prog[p], "len", H.len); set_f(prog[p], "time", H.time);
        break;
    case P_DEAD_COUNT: case P_DEAD_SCAN: case P_DEAD_SCATTER: case P_AABB_COLLAPSE:
    case P_BUCKET_SCAN: case P_GLOBAL_SCAN:
        set_u(prog[p], "numWorkGroups", G);
        break;
    case P_HIST: case P_SCATTER:
        set_u(prog[p], "passShift", u_pass_shift);
        set_u(prog[p], "numWorkGroups", G);
        break;
    case P_FORCE:
        set_f(prog[p], "theta", H.theta); set_f(prog[p], "dt", H.dt);
        set_f(prog[p], "elasticity", H.elasticity); set_u(prog[p], "wrapAround", H.wrap);
        set_f(prog[p], "softening", H.softening);
        set_u(prog[p], "mergingCollisionOrNeither", H.merging);
        set_u(prog[p], "staticOrDynamic", H.dynamic);
        break;
    default: break;
    }
    for (int b = 0; b < B_COUNT; b++) pglBindBufferBase(GL_SHADER_STORAGE_BUFFER, b, buf[b]);
    pglDispatchCompute(groups, 1, 1);
    pglMemoryBarrier(GL_SHADER_STORAGE_BARRIER_BIT);
    pglUseProgram(0);
    check_gl(pdef[p]);
}

static void download(int b, void *dst, size_t off, size_t n) {
    pglBindBuffer(GL_SHADER_STORAGE_BUFFER, buf[b]);
    pglGetBufferSubData(GL_SHADER_STORAGE_BUFFER, off, n, dst);
    pglBindBuffer(GL_SHADER_STORAGE_BUFFER, 0);
}
static void upload(int b, const void *src, size_t off, size_t n) {
    pglBindBuffer(GL_SHADER_STORAGE_BUFFER, buf[b]);
    pglBufferSubData(GL_SHADER_STORAGE_BUFFER, off, n, src);
    pglBindBuffer(GL_SHADER_STORAGE_BUFFER, 0);
}

static void dump(int step, const char *stage, int b) {
    char path[4096];
    snprintf(path, sizeof path, "%s/s%d_%s_%s.bin", outdir, step, stage, bname[b]);
    void *tmp = malloc(bsize[b]);
    download(b, tmp, 0, bsize[b]);
    FILE *f = fopen(path, "wb");
    if (!f) { fprintf(stderr, "cannot write %s\n", path); exit(2); }
    fwrite(tmp, 1, bsize[b], f);
    fclose(f);
    free(tmp);
}

/* per-stage timing, glFinish-bracketed as in simulation/BarnesHut.java:193-212 */
enum { T_RESET, T_DEAD, T_DECR, T_AABB, T_ENCODE, T_RADIX, T_BUILD, T_FILL, T_FORCE, T_MERGE, T_COUNT };
static const char *tname[T_COUNT] = {"reset", "dead", "decrement", "aabb", "encode", "radix", "build", "fill", "force", "merge"};
static double tsum[T_COUNT];
static double t0;
static void tbegin(void) { if (H.flags & 4) { pglFinish(); t0 = now_ms(); } }
static void tend(int k) { if (H.flags & 4) { pglFinish(); tsum[k] += now_ms() - t0; } }

static int dumping(int step) { return (H.flags & 1) && step >= (int)H.dump_from; }

/* Acceleration extraction: FORCE on the step's own state with velocities
 * zeroed and dt = 1 writes dstB.vel = 0 + a*1 = a exactly. All state the FORCE
 * kernel mutates is saved first and restored afterwards. */
static void extract_accel(int step) {
    size_t nb = bsize[B_BODIES_IN];
    char *src = malloc(nb), *dst = malloc(nb), *sim = malloc(bsize[B_SIM]), *mq = malloc(bsize[B_MERGE]);
    download(B_BODIES_IN, src, 0, nb);
    download(B_BODIES_OUT, dst, 0, nb);
    download(B_SIM, sim, 0, bsize[B_SIM]);
    download(B_MERGE, mq, 0, bsize[B_MERGE]);
    float *zv = malloc(nb);
    memcpy(zv, src, nb);
    for (uint32_t i = 0; i < N; i++) { zv[i * 8 + 4] = 0.f; zv[i * 8 + 5] = 0.f; zv[i * 8 + 6] = 0.f; }
    upload(B_BODIES_IN, zv, 0, nb);
    float saved_dt = H.dt;
    H.dt = 1.0f;
    dispatch(P_FORCE, G);
    H.dt = saved_dt;
    dump(step, "accel", B_BODIES_OUT);
    dump(step, "accel", B_SIM);
    upload(B_BODIES_IN, src, 0, nb);
    upload(B_BODIES_OUT, dst, 0, nb);
    upload(B_SIM, sim, 0, bsize[B_SIM]);
    upload(B_MERGE, mq, 0, bsize[B_MERGE]);
    free(src); free(dst); free(sim); free(mq); free(zv);
}

/* simulation/BarnesHut.java:99-164 */
static void step(int s) {
    int d = dumping(s);
    tbegin();
    u_reset_or_decrement = 1;                               /* resetValues() */
    dispatch(P_UPDATE, 1);
    tend(T_RESET);

    tbegin();                                               /* partitionDeadBodies() */
    dispatch(P_DEAD_COUNT, G);
    dispatch(P_DEAD_SCAN, 1);
    dispatch(P_DEAD_SCATTER, G);
    tend(T_DEAD);
    swap_names(B_MORTON_IN, B_MORTON_OUT);                  /* GPU.swapMortonAndIndexBuffers() */
    swap_names(B_INDEX_IN, B_INDEX_OUT);
    if (d) { dump(s, "dead", B_INDEX_IN); dump(s, "dead", B_SIM); dump(s, "dead", B_BODIES_IN); }

    tbegin();
    u_reset_or_decrement = 0;                               /* decrementDeadBodies() */
    dispatch(P_UPDATE, 1);
    tend(T_DECR);

    if (H.dynamic == 1) {                                   /* updateBounds() */
        tbegin();
        dispatch(P_AABB_REPOP, G);
        dispatch(P_AABB_COLLAPSE, 1);
        tend(T_AABB);
    }
    if (d) dump(s, "aabb", B_SIM);

    tbegin();
    dispatch(P_ENCODE, G);                                  /* generateMortonCodes() */
    tend(T_ENCODE);
    if (d) { dump(s, "encode", B_MORTON_IN); dump(s, "encode", B_INDEX_IN); }

    tbegin();                                               /* radixSort() */
    for (int pass = 0; pass < 16; pass++) {                 /* ceil(63/4) passes */
        u_pass_shift = (uint32_t)pass * 4u;
        dispatch(P_HIST, G);
        dispatch(P_BUCKET_SCAN, 16);
        dispatch(P_GLOBAL_SCAN, 16);
        dispatch(P_SCATTER, G);
        swap_names(B_MORTON_IN, B_MORTON_OUT);
        swap_names(B_INDEX_IN, B_INDEX_OUT);
    }
    tend(T_RADIX);
    if (d) { dump(s, "sort", B_MORTON_IN); dump(s, "sort", B_INDEX_IN); }

    tbegin();
    dispatch(P_TREE_BUILD, (N - 1 + 255) / 256);            /* buildBinaryRadixTree() */
    tend(T_BUILD);
    if (d) { dump(s, "build", B_INTERNAL); dump(s, "build", B_LEAF); }

    tbegin();                                               /* computeCOMAndLocation() */
    dispatch(P_INIT_LEAVES, G);
    for (int pass = 0; pass < 64; pass++) {
        double denom = 1.0;
        for (int k = 0; k < pass; k++) denom *= 2.0;
        int maxNodes = (int)((double)(N - 1) / denom);
        if (maxNodes < 4 * 256) maxNodes = 4 * 256;
        dispatch(P_PROPAGATE, (uint32_t)((maxNodes + 255) / 256));
        swap_names(B_WQ_IN, B_WQ_OUT);
    }
    tend(T_FILL);
    if (d) { dump(s, "fill", B_INTERNAL); dump(s, "fill", B_LEAF); }

    if (H.flags & 2) extract_accel(s);

    tbegin();
    dispatch(P_FORCE, G);                                   /* computeForce() */
    tend(T_FORCE);
    if (d) { dump(s, "force", B_BODIES_OUT); dump(s, "force", B_SIM); dump(s, "force", B_MERGE); }

    tbegin();
    dispatch(P_MERGE, G);                                   /* mergeBodies() */
    tend(T_MERGE);
    if (d) dump(s, "merge", B_BODIES_OUT);

    swap_names(B_BODIES_IN, B_BODIES_OUT);                  /* GPU.swapBodyBuffers() */
}

int main(int argc, char **argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: %s <shader_root> <input.bin> <outdir>\n", argv[0]);
        return 1;
    }
    shader_root = argv[1];
    outdir = argv[3];
    FILE *f = fopen(argv[2], "rb");
    if (!f) { fprintf(stderr, "cannot open %s\n", argv[2]); return 2; }
    if (fread(&H, sizeof H, 1, f) != 1 || H.magic != 0x4E494842u) { fprintf(stderr, "bad header\n"); return 2; }
    N = H.n;
    if (N < 2) { fprintf(stderr, "N must be >= 2\n"); return 2; }
    G = (N + 255) / 256;
    float *bodies = malloc((size_t)N * 32);
    if (fread(bodies, 32, N, f) != N) { fprintf(stderr, "short body data\n"); return 2; }
    fclose(f);

    create_context();
    GLint maxblock = 0;
    pglGetIntegerv(GL_MAX_SHADER_STORAGE_BLOCK_SIZE, &maxblock);
    if ((uint64_t)N * 80u > (uint64_t)(uint32_t)maxblock) {      /* gpu/GPU.java:201-209 */
        fprintf(stderr, "N=%u exceeds GL_MAX_SHADER_STORAGE_BLOCK_SIZE=%d\n", N, maxblock);
        return 5;
    }

    /* buffer sizes: gpu/GPU.java:199-341 */
    bsize[B_LEAF] = (size_t)N * 80;
    bsize[B_INTERNAL] = (size_t)(N - 1) * 80;
    bsize[B_SIM] = 8 * 4 + 16 * 4 + 100 * 4 + 100 * 4;
    bsize[B_BODIES_IN] = bsize[B_BODIES_OUT] = (size_t)N * 32;
    bsize[B_MORTON_IN] = bsize[B_MORTON_OUT] = (size_t)N * 8;
    bsize[B_INDEX_IN] = bsize[B_INDEX_OUT] = (size_t)N * 4;
    bsize[B_WQ_IN] = bsize[B_WQ_OUT] = (size_t)(4 + N) * 4;
    bsize[B_HIST] = (size_t)G * 16 * 4;
    bsize[B_SCANNED] = 4 + (size_t)G * 16 * 4 + 4;
    bsize[B_TOTALS] = 16 * 4 * 2;
    bsize[B_MERGE] = 2 * 4 + (size_t)N * 2 * 4;
    bsize[B_LOCKS] = (size_t)N * 4;
    pglGenBuffers(B_COUNT, buf);
    for (int b = 0; b < B_COUNT; b++) {
        void *z = calloc(1, bsize[b]);
        pglBindBuffer(GL_SHADER_STORAGE_BUFFER, buf[b]);
        pglBufferData(GL_SHADER_STORAGE_BUFFER, bsize[b], z, GL_DYNAMIC_COPY);
        free(z);
    }
    pglBindBuffer(GL_SHADER_STORAGE_BUFFER, 0);
    check_gl("buffers");

    /* sim packing: gpu/GPU.java:1096-1143 */
    uint32_t sim[224];
    memset(sim, 0, sizeof sim);
    sim[0] = N; sim[1] = N;
    memcpy(&sim[8], &H.bounds[0], 12);
    memcpy(&sim[12], &H.bounds[3], 12);
    float units[4] = {H.mass, H.density, H.len, H.time};
    memcpy(&sim[16], units, 16);
    upload(B_SIM, sim, 0, sizeof sim);
    upload(B_BODIES_IN, bodies, 0, (size_t)N * 32);

    double tc = now_ms();
    for (int p = 0; p < P_COUNT; p++) prog[p] = build_program(pdef[p]);
    double compile_ms = now_ms() - tc;

    /* report the std430 offsets GL assigned inside SimulationValues */
    const char *probe[] = {"SimulationValues.bounds.minCorner", "SimulationValues.bounds.maxCorner",
                           "SimulationValues.units.mass", "SimulationValues.units.gravitationalConstant",
                           "SimulationValues.uintDebug[0]", "SimulationValues.floatDebug[0]"};
    fprintf(stderr, "{\"sim_offsets\": {");
    for (int i = 0; i < 6; i++) {
        GLuint idx = pglGetProgramResourceIndex(prog[P_UPDATE], GL_BUFFER_VARIABLE, probe[i]);
        GLint off = -1; GLenum prop = GL_OFFSET;
        if (idx != GL_INVALID_INDEX) pglGetProgramResourceiv(prog[P_UPDATE], GL_BUFFER_VARIABLE, idx, 1, &prop, 1, NULL, &off);
        fprintf(stderr, "%s\"%s\": %d", i ? ", " : "", probe[i], off);
    }
    fprintf(stderr, "}}\n");

    dispatch(P_INIT, G);                                     /* BarnesHut.init() */
    pglFinish();

    double tstart = now_ms();
    for (uint32_t s = 0; s < H.steps; s++) {
        step((int)s);
        if (dumping((int)s) || s + 1 == H.steps) {
            dump((int)s, "end", B_BODIES_IN);
            dump((int)s, "end", B_SIM);
        }
    }
    pglFinish();
    double total = now_ms() - tstart;

    const char *threads = getenv("LP_NUM_THREADS");
    printf("{\"n\": %u, \"steps\": %u, \"compile_ms\": %.3f, \"total_ms\": %.3f, \"lp_num_threads\": \"%s\", \"renderer\": \"%s\", \"stage_ms\": {",
           N, H.steps, compile_ms, total, threads ? threads : "", (const char *)pglGetString(GL_RENDERER));
    for (int k = 0; k < T_COUNT; k++) printf("%s\"%s\": %.3f", k ? ", " : "", tname[k], tsum[k]);
    printf("}}\n");
    return 0;
}
