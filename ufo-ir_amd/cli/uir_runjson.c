/*
 * uir-runjson -- runs a ufo-core JSON graph (read -> ufo-ir tasks -> write) on
 * one GPU with libuir, as ufo-core's ufo-runjson runs it with the OpenCL
 * plugins.
 *
 *   uir-runjson [--device N] graph.json
 *
 * Relative paths in the graph are taken from the working directory.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "uir.h"

static char *slurp(const char *path)
{
    FILE *f = fopen(path, "rb");
    if (!f)
        return NULL;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *buf = n >= 0 ? (char *)malloc((size_t)n + 1) : NULL;
    if (buf && fread(buf, 1, (size_t)n, f) != (size_t)n) {
        free(buf);
        buf = NULL;
    }
    if (buf)
        buf[n] = 0;
    fclose(f);
    return buf;
}

int main(int argc, char **argv)
{
    int device = 0;
    const char *graph = NULL;
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "--device") && i + 1 < argc)
            device = atoi(argv[++i]);
        else if (argv[i][0] == '-') {
            fprintf(stderr, "usage: %s [--device N] graph.json\n", argv[0]);
            return 2;
        } else
            graph = argv[i];
    }
    if (!graph) {
        fprintf(stderr, "usage: %s [--device N] graph.json\n", argv[0]);
        return 2;
    }
    char *text = slurp(graph);
    if (!text) {
        fprintf(stderr, "%s: cannot read %s\n", argv[0], graph);
        return 1;
    }
    uir_json_report rep;
    memset(&rep, 0, sizeof rep);
    uir_status s = uir_json_run(text, NULL, device, &rep);
    free(text);
    if (s != UIR_OK) {
        fprintf(stderr, "%s: %s: %s\n", argv[0], uir_status_string(s), uir_last_error());
        return 1;
    }
    printf("%u frame(s) through %u task(s) in %.3f s\n", rep.frames, rep.tasks, rep.seconds);
    return 0;
}
