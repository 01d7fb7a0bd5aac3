// Shared error reporting of the host library (song_host.h).
#pragma once

int host_fail(int code, const char *fmt, ...);

// worker threads: `threads` if positive, else SONG_THREADS, else min(16, cores)
int host_threads(int threads);
