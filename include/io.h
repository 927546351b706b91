// io.h - drop-in for the reference's ingestion ring
// (SebastianBoteroLeonik/bzip2gpu include/io.h:17-66, src/utils/io.cpp).
// A reader thread fills up to n_buffers pinned host chunks of buffer_size
// bytes; fetch() hands them out in order with a sequential chunk_idx and
// returns -1 once the input is exhausted.  The caller releases each chunk
// with hipHostFree.  Differences from the reference, all on its failure
// paths: no zero-length chunk is ever handed out, an unreadable input ends
// the stream instead of blocking fetch() forever, a read error part-way
// through the input ends it with `failed` set (the chunks read before the
// error are still handed out), and the reader thread is joined by the
// destructor.
#ifndef BZG_IO_H
#define BZG_IO_H

#include <stdint.h>

#include <condition_variable>
#include <istream>
#include <mutex>
#include <string>
#include <thread>

typedef struct BZFileInputStream {
  std::istream *in_stream;
  bool owns_stream = false;
  int n_buffers;
  int buffer_size;
  int head_idx = 0;
  int tail_idx = -1;
  int len = 0;
  int chunks_read = 0;
  bool is_closed = false;
  bool failed = false;  // input could not be opened, or a read failed
  uint8_t **buffers = nullptr;
  int *buffer_lengths = nullptr;
  std::mutex mtx;
  std::condition_variable not_full_cond;
  std::condition_variable not_empty_cond;
  std::thread reader;

  BZFileInputStream(int _n_buffers, int _buffer_size);
  BZFileInputStream(int _n_buffers, int _buffer_size, std::string filename);
  // Extension (no counterpart in the reference): reads from a caller-owned
  // stream, which must outlive this object.
  BZFileInputStream(int _n_buffers, int _buffer_size, std::istream *stream);
  ~BZFileInputStream();

  void setup_buffers(int _n_buffers, int _buffer_size);
  void reader_thread_task();
  int fetch(uint8_t *&result, int &chunk_idx);
} BZFileInputStream;

#endif  // BZG_IO_H
