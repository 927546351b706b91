// io.cpp - BZFileInputStream (reference: src/utils/io.cpp:2-55).
#include "io.h"

#include <hip/hip_runtime.h>

#include <fstream>
#include <iostream>

BZFileInputStream::BZFileInputStream(int _n_buffers, int _buffer_size) {
  in_stream = &std::cin;
  setup_buffers(_n_buffers, _buffer_size);
}

BZFileInputStream::BZFileInputStream(int _n_buffers, int _buffer_size, std::string filename) {
  in_stream = new std::ifstream(filename, std::ios::in | std::ios::binary);
  owns_stream = true;
  setup_buffers(_n_buffers, _buffer_size);
}

BZFileInputStream::BZFileInputStream(int _n_buffers, int _buffer_size, std::istream *stream) {
  in_stream = stream;
  setup_buffers(_n_buffers, _buffer_size);
}

BZFileInputStream::~BZFileInputStream() {
  {
    std::lock_guard<std::mutex> lk(mtx);
    is_closed = true;
  }
  not_full_cond.notify_all();
  if (reader.joinable()) reader.join();
  // chunks never fetched
  for (int i = 0; i < n_buffers; i++)
    if (buffers[i]) (void)hipHostFree(buffers[i]);
  delete[] buffers;
  delete[] buffer_lengths;
  if (owns_stream) delete in_stream;
}

void BZFileInputStream::setup_buffers(int _n_buffers, int _buffer_size) {
  n_buffers = _n_buffers < 1 ? 1 : _n_buffers;
  buffer_size = _buffer_size;
  buffers = new uint8_t *[n_buffers];
  buffer_lengths = new int[n_buffers];
  for (int i = 0; i < n_buffers; i++) {
    buffers[i] = nullptr;
    buffer_lengths[i] = -1;
  }
  if (!in_stream || !*in_stream || buffer_size <= 0) {
    failed = !in_stream || !*in_stream;
    is_closed = true;
    return;
  }
  reader = std::thread(&BZFileInputStream::reader_thread_task, this);
}

void BZFileInputStream::reader_thread_task() {
  for (;;) {
    std::unique_lock<std::mutex> lock(mtx);
    not_full_cond.wait(lock, [this] { return len < n_buffers || is_closed; });
    if (is_closed) break;
    lock.unlock();
    uint8_t *chunk = nullptr;
    if (hipHostMalloc((void **)&chunk, (size_t)buffer_size, hipHostMallocDefault) != hipSuccess) {
      lock.lock();
      failed = true;
      is_closed = true;
      not_empty_cond.notify_all();
      break;
    }
    in_stream->read(reinterpret_cast<char *>(chunk), buffer_size);
    int got = (int)in_stream->gcount();
    // a short read ends the input: end of file (failbit + eofbit) or a read
    // error (badbit), which is reported through `failed`
    bool eof = !*in_stream;
    bool bad = in_stream->bad();
    lock.lock();
    if (bad) failed = true;
    if (got > 0) {
      tail_idx = (tail_idx + 1) % n_buffers;
      buffers[tail_idx] = chunk;
      buffer_lengths[tail_idx] = got;
      len++;
    } else {
      (void)hipHostFree(chunk);
    }
    if (eof) is_closed = true;
    not_empty_cond.notify_all();
    if (eof) break;
  }
}

int BZFileInputStream::fetch(uint8_t *&result, int &chunk_idx) {
  std::unique_lock<std::mutex> lock(mtx);
  not_empty_cond.wait(lock, [this] { return len > 0 || is_closed; });
  if (len == 0) return -1;
  result = buffers[head_idx];
  int res_len = buffer_lengths[head_idx];
  buffers[head_idx] = nullptr;
  buffer_lengths[head_idx] = -1;
  head_idx = (head_idx + 1) % n_buffers;
  len--;
  chunk_idx = chunks_read++;
  not_full_cond.notify_all();
  return res_len;
}
