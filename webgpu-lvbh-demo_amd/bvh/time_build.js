'use strict'
// Wall time of initBVHBuild(device).build(TRI_ARRAY, NUM_TRIS, MODEL_BOUNDS)
// measured the way main.js:87-90 does (a clock around the awaited build):
// the triangles' upload, the output buffers' allocation and the build.
//   node time_build.js <tri.f32> <n> <bounds.json|-> [reps]
// TRI_ARRAY is the file's Float32Array, used without a copy; main.js passes
// a JS Array, whose conversion to float32 is the caller's cost in the
// reference as well.  Prints one JSON line: every repetition's milliseconds
// (the first includes the context's one-time set-up) and the best.
const fs = require('fs')
const bvh = require(__dirname)

async function main () {
  const [triPath, nStr, boundsArg, repsArg] = process.argv.slice(2)
  const n = parseInt(nStr, 10)
  const reps = parseInt(repsArg || '4', 10)
  let raw = fs.readFileSync(triPath)
  if (raw.byteOffset % 4 !== 0) raw = Buffer.from(raw)
  if (raw.length !== 48 * n) throw new Error(`${triPath}: ${raw.length} bytes, expected ${48 * n}`)
  const TRI_ARRAY = new Float32Array(raw.buffer, raw.byteOffset, raw.length / 4)
  const MODEL_BOUNDS = boundsArg === '-' ? null : JSON.parse(fs.readFileSync(boundsArg, 'utf8'))
  const device = bvh.createDevice({ ordinal: 0 })
  const BVH = bvh.initBVHBuild(device)
  const ms = []
  for (let r = 0; r < reps; r++) {
    const t0 = process.hrtime.bigint()
    const { BVH_BUFFER, O_TRIANGLE_BUFFER } = await BVH.build(TRI_ARRAY, n, MODEL_BOUNDS)
    const t1 = process.hrtime.bigint()
    ms.push(Number(t1 - t0) / 1e6)
    BVH_BUFFER.destroy()
    O_TRIANGLE_BUFFER.destroy()
  }
  console.log(JSON.stringify({ n, ms, best_ms: Math.min(...ms.slice(1)) }))
}

main().catch((e) => { console.error(e); process.exit(1) })
